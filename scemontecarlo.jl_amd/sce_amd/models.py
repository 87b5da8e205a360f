"""Synthetic SCE model generator — the stand-in for SCEFitting's ``multipole_terms``.

The reference tiles a fitted training-cell model (``MultipoleTerm`` lists from
SCEFitting, src/hamiltonian.jl:267-400).  SCEFitting is not part of the reference
tree, so this module writes seeded ``MultipoleTerm`` lists for cubic lattices:

* lattices: simple cubic (1 atom), bcc (2-atom cubic cell), fcc (4-atom cubic cell);
* one term per physical cluster (SCEFitting's canonical convention), clusters
  enumerated up to lattice translation with the first member in the home cell
  (``shifts[0] == 0``, the reference's anchoring rule, hamiltonian.jl:306);
* body-2 bonds on the first k neighbour shells, one template per (l1, l2) block;
* bcc body-3 triangles (1NN, 1NN, 2NN) and body-4 tetrahedra (four 1NN and two
  2NN edges) — the bcc 1NN graph is bipartite, so it has no all-1NN triangle;
* coefficients: raw ``coef ~ N(0, scale)`` per template (scale = 1 meV by default),
  and either dense random ``folded`` tensors (entries N(0,1)/sqrt(size)) or the
  isotropic SALC tensors (identity for (l,l) pairs, Levi-Civita for l=1 triangles,
  the three pairings for l=1 tetrahedra).

Hand-built fixtures of the reference's unit suite (test/unit/fixtures.jl:94-131)
are reproduced value for value in :func:`chain_terms` etc.
"""
from __future__ import annotations

import itertools
import math
from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

import numpy as np


@dataclass
class MultipoleTerm:
    """One raw (unscaled) SCE term: ``coef``, member ``atoms`` (1-based training-cell
    atoms), per-member integer lattice ``shifts`` (``shifts[0] == (0,0,0)``), angular
    momenta ``ls`` and the rank-``body`` tensor ``folded`` of shape ``(2l+1, ...)``."""

    coef: float
    atoms: List[int]
    shifts: List[Tuple[int, int, int]]
    ls: List[int]
    folded: np.ndarray = field(repr=False)

    @property
    def body(self) -> int:
        return len(self.atoms)


BASIS = {
    "sc": np.array([[0.0, 0.0, 0.0]]),
    "bcc": np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.5]]),
    "fcc": np.array([[0.0, 0.0, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]]),
}

FOURPI = 4.0 * math.pi
N1 = math.sqrt(3.0 / FOURPI)          # Z_1m = N1 * (y, z, x)[m + 2]


# ---------------------------------------------------------------------------
# reference unit-suite fixtures (test/unit/fixtures.jl:94-131, test_hamiltonian.jl)
# ---------------------------------------------------------------------------

def chain_terms(J: float) -> List[MultipoleTerm]:
    """Self-image l=1 chain along x (fixtures.jl:94-103): both directed members."""
    f = np.zeros((3, 3))
    f[0, 0] = f[1, 1] = f[2, 2] = 1.0
    raw = J / (2 * N1 ** 2 * FOURPI)
    z, x, mx = (0, 0, 0), (1, 0, 0), (-1, 0, 0)
    return [MultipoleTerm(raw, [1, 1], [z, x], [1, 1], f.copy()),
            MultipoleTerm(raw, [1, 1], [z, mx], [1, 1], f.copy())]


def threebody_terms(J: float) -> List[MultipoleTerm]:
    """3-body chain (0, +x, +2x) — triplet fast path (fixtures.jl:109-117)."""
    f = np.zeros((3, 3, 3))
    f[1, 1, 1] = 1.0
    f[0, 2, 0] = 0.7
    f[2, 0, 1] = -0.4
    return [MultipoleTerm(J, [1, 1, 1], [(0, 0, 0), (1, 0, 0), (2, 0, 0)], [1, 1, 1], f)]


def fourbody_terms(J: float) -> List[MultipoleTerm]:
    """4-body chain (0, +x, +2x, +3x) — general branch (fixtures.jl:122-131)."""
    f = np.zeros((3, 3, 3, 3))
    f[1, 1, 1, 1] = 1.0
    f[0, 2, 1, 0] = 0.6
    f[2, 1, 0, 2] = -0.3
    return [MultipoleTerm(J, [1, 1, 1, 1], [(0, 0, 0), (1, 0, 0), (2, 0, 0), (3, 0, 0)],
                          [1, 1, 1, 1], f)]


def field_term(c0: float) -> List[MultipoleTerm]:
    """Single-site l=1 term V(e) = c0 N1 e_z (test_hamiltonian.jl:110-122)."""
    f = np.zeros(3)
    f[1] = 1.0
    return [MultipoleTerm(c0 / math.sqrt(FOURPI), [1], [(0, 0, 0)], [1], f)]


def dimer_terms(J: float) -> List[MultipoleTerm]:
    """An isotropic l=1 dimer on atoms 1-2 of a 4-atom cell (atoms 3-4 inactive):
    E = J e1.e2 — the hand-built analogue of the reference's SCEFitting dimer
    (fixtures.jl:20-31)."""
    f = np.eye(3)
    raw = J / (N1 ** 2 * FOURPI)
    return [MultipoleTerm(raw, [1, 2], [(0, 0, 0), (0, 0, 0)], [1, 1], f)]


# ---------------------------------------------------------------------------
# lattice clusters
# ---------------------------------------------------------------------------

def _sites_within(lattice: str, rmax: float):
    basis = BASIS[lattice]
    R = int(math.ceil(rmax)) + 1
    out = []
    for c in itertools.product(range(-R, R + 1), repeat=3):
        for a, b in enumerate(basis):
            out.append((a + 1, tuple(int(x) for x in c), np.asarray(c, float) + b))
    return out


def shell_distances(lattice: str, nshells: int) -> List[float]:
    basis = BASIS[lattice]
    ds = set()
    for a, b in enumerate(basis):
        for (_, _, p) in _sites_within(lattice, 3.0):
            d = float(np.linalg.norm(p - b))
            if d > 1e-9:
                ds.add(round(d, 9))
    return sorted(ds)[:nshells]


def _canon(members):
    """Translation-canonical form of a cluster given as (atom, cell) members:
    the lexicographically smallest sorted tuple over translations that put one
    member in the home cell."""
    best = None
    for (_, c0) in members:
        tr = tuple(sorted((a, (c[0] - c0[0], c[1] - c0[1], c[2] - c0[2])) for (a, c) in members))
        # reorder so that a member in the home cell comes first
        home = [m for m in tr if m[1] == (0, 0, 0)]
        rest = [m for m in tr if m[1] != (0, 0, 0)]
        key = tuple(home + rest)
        if best is None or key < best:
            best = key
    return best


def _pos(lattice, a, c):
    return np.asarray(c, float) + BASIS[lattice][a - 1]


def bonds(lattice: str, shell: int) -> List[Tuple[Tuple[int, Tuple[int, int, int]], ...]]:
    """Unique bonds (up to translation) of neighbour shell ``shell`` (1-based)."""
    d = shell_distances(lattice, shell)[shell - 1]
    basis = BASIS[lattice]
    seen = set()
    out = []
    for a0 in range(1, len(basis) + 1):
        p0 = _pos(lattice, a0, (0, 0, 0))
        for (a, c, p) in _sites_within(lattice, d + 1):
            if abs(np.linalg.norm(p - p0) - d) < 1e-9:
                key = _canon([(a0, (0, 0, 0)), (a, c)])
                if key not in seen:
                    seen.add(key)
                    out.append(key)
    return sorted(out)


def _clusters_by_edges(lattice: str, body: int, want: Sequence[float]):
    """Clusters of ``body`` sites whose sorted pair distances equal ``want``."""
    dmax = max(want)
    basis = BASIS[lattice]
    near = [(a, c, p) for (a, c, p) in _sites_within(lattice, dmax + 1)]
    want = sorted(round(x, 9) for x in want)
    seen = set()
    out = []
    for a0 in range(1, len(basis) + 1):
        p0 = _pos(lattice, a0, (0, 0, 0))
        cand = [(a, c, p) for (a, c, p) in near
                if 1e-9 < np.linalg.norm(p - p0) < dmax + 1e-9]
        for combo in itertools.combinations(cand, body - 1):
            pts = [p0] + [x[2] for x in combo]
            ds = sorted(round(float(np.linalg.norm(pts[i] - pts[j])), 9)
                        for i in range(body) for j in range(i + 1, body))
            if ds != want:
                continue
            key = _canon([(a0, (0, 0, 0))] + [(x[0], x[1]) for x in combo])
            if key not in seen:
                seen.add(key)
                out.append(key)
    return sorted(out)


def bcc_triangles():
    d1, d2 = shell_distances("bcc", 2)
    return _clusters_by_edges("bcc", 3, [d1, d1, d2])


def bcc_tetrahedra():
    d1, d2 = shell_distances("bcc", 2)
    return _clusters_by_edges("bcc", 4, [d1, d1, d1, d1, d2, d2])


# ---------------------------------------------------------------------------
# tensors
# ---------------------------------------------------------------------------

def _levi_civita():
    f = np.zeros((3, 3, 3))
    for (i, j, k), s in (((0, 1, 2), 1), ((1, 2, 0), 1), ((2, 0, 1), 1),
                         ((0, 2, 1), -1), ((2, 1, 0), -1), ((1, 0, 2), -1)):
        f[i, j, k] = s
    return f


def _three_pairings():
    f = np.zeros((3, 3, 3, 3))
    I = np.eye(3)
    for a, b, c, d in itertools.product(range(3), repeat=4):
        f[a, b, c, d] = I[a, b] * I[c, d] + I[a, c] * I[b, d] + I[a, d] * I[b, c]
    return f


def _tensor(rng, ls, isotropic):
    shape = tuple(2 * l + 1 for l in ls)
    if not isotropic:
        t = rng.standard_normal(shape)
        return t / math.sqrt(t.size)
    if len(ls) == 2:
        return np.eye(shape[0]) if ls[0] == ls[1] else None
    if len(ls) == 3 and all(l == 1 for l in ls):
        return _levi_civita()
    if len(ls) == 4 and all(l == 1 for l in ls):
        return _three_pairings()
    if len(ls) == 1:
        return None
    raise ValueError(f"no isotropic tensor for ls={ls}")


def _term(members, coef, ls, folded):
    atoms = [a for (a, _) in members]
    shifts = [c for (_, c) in members]
    return MultipoleTerm(float(coef), atoms, shifts, list(ls), folded)


def lattice_model(lattice: str, *, pair_shells: int = 1, pair_lmax: int = 1,
                  triangles: bool = False, tetrahedra: bool = False,
                  onsite_l2: bool = False, isotropic: bool = False,
                  seed: int = 0, scale: float = 1e-3) -> Tuple[int, List[MultipoleTerm]]:
    """Seeded synthetic model on a cubic lattice: returns ``(n_cell_atoms, terms)``.

    Term order: on-site body-1 l=2 anisotropy, then bonds shell by shell with the
    (l1, l2) templates (1,1), (1,2), (2,1), (2,2) as allowed by ``pair_lmax``,
    then bcc triangles, then bcc tetrahedra.
    """
    rng = np.random.default_rng(seed)
    na = len(BASIS[lattice])
    terms: List[MultipoleTerm] = []
    if onsite_l2:
        for a in range(1, na + 1):
            f = rng.standard_normal(5) / math.sqrt(5)
            terms.append(MultipoleTerm(float(rng.normal(0, scale)), [a], [(0, 0, 0)], [2], f))
    lpairs = [(l1, l2) for l1 in range(1, pair_lmax + 1) for l2 in range(1, pair_lmax + 1)]
    for sh in range(1, pair_shells + 1):
        for b in bonds(lattice, sh):
            for (l1, l2) in lpairs:
                f = _tensor(rng, (l1, l2), isotropic)
                c = rng.normal(0, scale)
                if f is None:
                    continue
                terms.append(_term(b, c, (l1, l2), f))
    if triangles:
        if lattice != "bcc":
            raise ValueError("triangles are defined for bcc only")
        for tri in bcc_triangles():
            f = _tensor(rng, (1, 1, 1), isotropic)
            terms.append(_term(tri, rng.normal(0, scale), (1, 1, 1), f))
    if tetrahedra:
        if lattice != "bcc":
            raise ValueError("tetrahedra are defined for bcc only")
        for tet in bcc_tetrahedra():
            f = _tensor(rng, (1, 1, 1, 1), isotropic)
            terms.append(_term(tet, rng.normal(0, scale), (1, 1, 1, 1), f))
    return na, terms


# ---------------------------------------------------------------------------
# the workload configurations
# ---------------------------------------------------------------------------

def config_model(name: str, seed: int = 0):
    """``(n_cell_atoms, terms, dims)`` of a named workload.

    * ``sc8``     simple cubic 8^3 (512 spins): body-1 l=2 + body-2 l=1 on 2 shells
    * ``bcc16``   bcc 16^3 (8192 spins): body-2 l<=2 to 3rd shell + l=1 triangles
    * ``fcc12``   fcc 12^3 (6912 spins): body-2 l<=2 to 4th shell
    * ``bcc32``   bcc 32^3 (65536 spins): body-2 l<=2 to 3rd shell + l=1 triangles
                  + l=1 tetrahedra, isotropic SALC tensors
    * ``bcc32d``  the same with dense random tensors
    """
    if name == "sc8":
        na, t = lattice_model("sc", pair_shells=2, pair_lmax=1, onsite_l2=True, seed=seed)
        return na, t, (8, 8, 8)
    if name == "bcc16":
        na, t = lattice_model("bcc", pair_shells=3, pair_lmax=2, triangles=True, seed=seed)
        return na, t, (16, 16, 16)
    if name == "fcc12":
        na, t = lattice_model("fcc", pair_shells=4, pair_lmax=2, seed=seed)
        return na, t, (12, 12, 12)
    if name == "bcc32":
        na, t = lattice_model("bcc", pair_shells=3, pair_lmax=2, triangles=True,
                              tetrahedra=True, isotropic=True, seed=seed)
        return na, t, (32, 32, 32)
    if name == "bcc32d":
        na, t = lattice_model("bcc", pair_shells=3, pair_lmax=2, triangles=True,
                              tetrahedra=True, isotropic=False, seed=seed)
        return na, t, (32, 32, 32)
    raise KeyError(name)


def random_config(n_sites: int, rng) -> np.ndarray:
    """Uniform unit spins (normalised Gaussian triples, state.jl:80-86), n x 3."""
    v = rng.standard_normal((n_sites, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


# ---------------------------------------------------------------------------
# setup files and general cells (the model-input side, src/hamiltonian.jl:396-400)
# ---------------------------------------------------------------------------

@dataclass
class Setup:
    """A cell and its interaction defaults, read from a setup file in the layout the
    reference's bench fixtures read (bench/fixtures.jl:68, bench/assets/nd2fe14b.toml):
    ``[structure]`` lattice vectors as rows (Å), fractional ``positions``, 1-based
    ``species`` per atom and ``species_labels``; ``[interaction]`` ``nbody``,
    ``cutoff`` (Å, ``inf`` allowed), per-species ``lmax`` and ``isotropy``;
    ``[symmetry]`` ``tol``."""

    lattice: np.ndarray
    positions: np.ndarray
    species: List[int]
    species_labels: List[str]
    nbody: int = 2
    cutoff: float = math.inf
    lmax: List[int] = field(default_factory=list)
    isotropy: bool = True
    tol: float = 1e-5

    @property
    def n_atoms(self) -> int:
        return len(self.species)

    def cartesian(self) -> np.ndarray:
        return self.positions @ self.lattice


def read_setup(path: str) -> Setup:
    """Parse a setup file (TOML) into a :class:`Setup`; raises ``ValueError`` on a
    malformed cell or interaction section."""
    import tomli
    with open(path, "rb") as f:
        d = tomli.load(f)
    if "structure" not in d:
        raise ValueError(f"{path}: no [structure] section")
    s, it = d["structure"], d.get("interaction", {})
    lat = np.asarray(s["lattice"], np.float64)
    pos = np.asarray(s["positions"], np.float64)
    sp = [int(x) for x in s["species"]]
    labels = [str(x) for x in s.get("species_labels", [str(k) for k in range(1, max(sp, default=0) + 1)])]
    if lat.shape != (3, 3) or abs(np.linalg.det(lat)) < 1e-12:
        raise ValueError(f"{path}: lattice must be three independent 3-vectors")
    if pos.ndim != 2 or pos.shape[1] != 3 or len(sp) != pos.shape[0] or not sp:
        raise ValueError(f"{path}: positions ({pos.shape}) and species ({len(sp)}) disagree")
    if min(sp) < 1 or max(sp) > len(labels):
        raise ValueError(f"{path}: species must index species_labels (1..{len(labels)})")
    lmax = [int(x) for x in it.get("lmax", [1] * len(labels))]
    if len(lmax) != len(labels) or min(lmax) < 0:
        raise ValueError(f"{path}: lmax needs one entry >= 0 per species ({len(labels)})")
    st = Setup(lat, pos, sp, labels, int(it.get("nbody", 2)), float(it.get("cutoff", math.inf)), lmax,
               bool(it.get("isotropy", True)), float(d.get("symmetry", {}).get("tol", 1e-5)))
    _check_interaction(st.nbody, st.cutoff)
    return st


def _check_interaction(nbody, cutoff):
    if nbody not in (1, 2, 3):
        raise ValueError(f"nbody must be 1, 2 or 3; got {nbody}")
    if not cutoff > 0:
        raise ValueError(f"cutoff must be > 0; got {cutoff}")
    if nbody == 3 and math.isinf(cutoff):
        raise ValueError("triplet clusters need a finite cutoff")


def setup_magmoms(setup: Setup, m: float = 2.0) -> List[float]:
    """Per-atom moments (μ_B): ``m`` on atoms of species with ``lmax >= 1``, 0 on
    the others (bench/fixtures.jl:80-88 puts 0 on boron, the Nd₂Fe₁₄B cell's only
    ``lmax = 0`` species)."""
    return [float(m) if setup.lmax[s - 1] >= 1 else 0.0 for s in setup.species]


def _image_range(setup: Setup, r: float):
    """Integer shifts per axis that cover every image within ``r`` of the cell."""
    recip = np.linalg.inv(setup.lattice).T
    n = [int(math.ceil(r * float(np.linalg.norm(recip[i])))) + 1 for i in range(3)]
    return itertools.product(*(range(-k, k + 1) for k in n))


def _neighbours(setup: Setup, a0: int, r: float, magnetic):
    """(atom, shift, position) of every magnetic site within ``r`` of home atom ``a0``
    (1-based), the site itself excluded."""
    P = setup.cartesian()
    out = []
    for c in _image_range(setup, r):
        off = np.asarray(c, np.float64) @ setup.lattice
        for a in range(setup.n_atoms):
            if not magnetic[a]:
                continue
            p = P[a] + off
            d = float(np.linalg.norm(p - P[a0 - 1]))
            if 1e-9 < d <= r + setup.tol:
                out.append((a + 1, tuple(int(x) for x in c), p))
    return out


def setup_pairs(setup: Setup, cutoff: float, magnetic):
    """Translation-unique pair clusters between magnetic atoms: every pair within
    ``cutoff`` (Å) — an atom with its own periodic images included — or, for
    ``cutoff = inf``, one pair per two distinct atoms of the cell at its minimum
    image (ties resolved to the smallest shift tuple; an atom is not paired with
    its own images, which have no single minimum).  The minimum image is searched
    over every shift within the cell's diameter, so skewed, unreduced cells
    resolve it as reduced ones do."""
    seen, out = set(), []
    if math.isinf(cutoff):
        P = setup.cartesian()
        diameter = float(sum(np.linalg.norm(v) for v in setup.lattice))
        shifts = sorted(_image_range(setup, diameter))
        for a0 in range(1, setup.n_atoms + 1):
            for a in range(a0 + 1, setup.n_atoms + 1):
                if not (magnetic[a0 - 1] and magnetic[a - 1]):
                    continue
                d = [float(np.linalg.norm(P[a - 1] + np.asarray(c, np.float64) @ setup.lattice - P[a0 - 1]))
                     for c in shifts]
                best = min(d)
                c = next(s for s, x in zip(shifts, d) if x <= best + setup.tol)
                key = _canon([(a0, (0, 0, 0)), (a, c)])
                if key not in seen:
                    seen.add(key)
                    out.append(key)
        return sorted(out)
    for a0 in range(1, setup.n_atoms + 1):
        if not magnetic[a0 - 1]:
            continue
        for (a, c, _) in _neighbours(setup, a0, cutoff, magnetic):
            key = _canon([(a0, (0, 0, 0)), (a, c)])
            if key not in seen:
                seen.add(key)
                out.append(key)
    return sorted(out)


def setup_triplets(setup: Setup, cutoff: float, magnetic):
    """Translation-unique triplet clusters of magnetic atoms whose three pair
    distances are all within ``cutoff`` (Å)."""
    seen, out = set(), []
    for a0 in range(1, setup.n_atoms + 1):
        if not magnetic[a0 - 1]:
            continue
        nb = _neighbours(setup, a0, cutoff, magnetic)
        for x, y in itertools.combinations(nb, 2):
            if float(np.linalg.norm(x[2] - y[2])) > cutoff + setup.tol:
                continue
            key = _canon([(a0, (0, 0, 0)), (x[0], x[1]), (y[0], y[1])])
            if key not in seen:
                seen.add(key)
                out.append(key)
    return sorted(out)


_GAUNT = {}


def real_gaunt(ls: Sequence[int]) -> np.ndarray:
    """Rotation-invariant coupling of real harmonics: the tensor
    ∫ Z_{l1 m1} Z_{l2 m2} Z_{l3 m3} dΩ over the sphere, scaled to unit Frobenius
    norm (zeros when the l's admit no invariant).  The integral is exact: a
    Gauss–Legendre rule in cos θ and a uniform rule in φ, both of degree above
    l1 + l2 + l3, over the library's own Z_lm rows."""
    key = tuple(int(l) for l in ls)
    if key in _GAUNT:
        return _GAUNT[key]
    from . import _lib as L
    lmax = max(key)
    x, w = np.polynomial.legendre.leggauss(lmax * 3 // 2 + 4)
    nphi = 3 * lmax + 4
    phi = 2.0 * math.pi * np.arange(nphi) / nphi
    st = np.sqrt(1.0 - x * x)
    dirs = np.stack([np.outer(st, np.cos(phi)), np.outer(st, np.sin(phi)),
                     np.repeat(x[:, None], nphi, axis=1)], axis=-1).reshape(-1, 3)
    wt = np.repeat(w, nphi) * (2.0 * math.pi / nphi)
    z = np.empty((len(dirs), (lmax + 1) ** 2))
    L.check(L.lib().sce_host_zrows(lmax, len(dirs), L.ptr(np.ascontiguousarray(dirs)), L.ptr(z)))
    blk = [z[:, l * l:(l + 1) * (l + 1)] for l in key]
    t = np.einsum("p,pa,pb,pc->abc", wt, *blk)
    nrm = float(np.linalg.norm(t))
    t = t / nrm if nrm > 1e-12 else np.zeros_like(t)
    _GAUNT[key] = t
    return t


def _pair_templates(la, lb, isotropic):
    if isotropic:
        return [(l, l) for l in range(1, min(la, lb) + 1)]
    return [(l1, l2) for l1 in range(1, la + 1) for l2 in range(1, lb + 1) if (l1 + l2) % 2 == 0]


def _triplet_templates(la, lb, lc):
    return [(l1, l2, l3) for l1 in range(1, la + 1) for l2 in range(1, lb + 1) for l3 in range(1, lc + 1)
            if (l1 + l2 + l3) % 2 == 0 and abs(l1 - l2) <= l3 <= l1 + l2]


def setup_model(setup: Setup, *, nbody: int = None, cutoff: float = None, lmax: Sequence[int] = None,
                isotropy: bool = None, seed: int = 13, scale: float = 0.01) -> Tuple[int, List[MultipoleTerm]]:
    """Seeded synthetic model on a setup file's cell: ``(n_cell_atoms, terms)``, the
    stand-in for ``multipole_terms`` of an SCEFitting model built on the same cell
    (src/hamiltonian.jl:396-400; bench/fixtures.jl:66-74 draws its coefficients as
    ``0.01·randn``).  Keywords left ``None`` take the file's ``[interaction]`` values.

    Atoms of species with ``lmax = 0`` carry no term.  Term order: on-site
    (anisotropic models only: even l = 2, 4, … ≤ the species' lmax, per atom),
    then pairs in :func:`setup_pairs` order, then triplets in
    :func:`setup_triplets` order; each cluster lists its templates in
    lexicographic ``ls`` order, and every template draws ``coef ~ N(0, scale)``.
    Templates keep the total l even (time-reversal even).  Isotropic tensors are
    the identity on (l, l) pairs and :func:`real_gaunt` on triplets (the
    triangle-rule l's); anisotropic ones are dense N(0, 1)/sqrt(size).  This is
    not SCEFitting's symmetry-adapted basis (the site-symmetry SALCs need the
    space group): term counts and coefficients differ from a fitted model's,
    the tiling, walk and contraction shapes do not."""
    nbody = setup.nbody if nbody is None else int(nbody)
    cutoff = setup.cutoff if cutoff is None else float(cutoff)
    lm = list(setup.lmax if lmax is None else [int(x) for x in lmax])
    iso = setup.isotropy if isotropy is None else bool(isotropy)
    _check_interaction(nbody, cutoff)
    if len(lm) != len(setup.species_labels):
        raise ValueError(f"lmax needs one entry per species ({len(setup.species_labels)})")
    la = [lm[s - 1] for s in setup.species]
    magnetic = [l >= 1 for l in la]
    rng = np.random.default_rng(seed)
    terms: List[MultipoleTerm] = []
    if not iso:
        for a in range(1, setup.n_atoms + 1):
            for l in range(2, la[a - 1] + 1, 2):
                f = rng.standard_normal(2 * l + 1) / math.sqrt(2 * l + 1)
                terms.append(MultipoleTerm(float(rng.normal(0, scale)), [a], [(0, 0, 0)], [l], f))
    if nbody >= 2:
        for key in setup_pairs(setup, cutoff, magnetic):
            for ls in _pair_templates(la[key[0][0] - 1], la[key[1][0] - 1], iso):
                f = np.eye(2 * ls[0] + 1) if iso else _tensor(rng, ls, False)
                terms.append(_term(key, rng.normal(0, scale), ls, f))
    if nbody >= 3:
        for key in setup_triplets(setup, cutoff, magnetic):
            for ls in _triplet_templates(*(la[m[0] - 1] for m in key)):
                f = real_gaunt(ls) if iso else _tensor(rng, ls, False)
                if iso and not np.any(f):
                    continue
                terms.append(_term(key, rng.normal(0, scale), ls, f))
    return setup.n_atoms, terms
