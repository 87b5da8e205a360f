"""Shared model fixtures of the test suite (the reference's test/unit/fixtures.jl
analogues, built with the hand-made term lists and the lattice generator)."""
import numpy as np

import oracle as O
import sce_amd as S
from sce_amd import models as M

ZB = (0.3, -1.2, 2.0)              # tesla, test_gpu.jl:23
ZMM_DIMER = [2.2, 2.2, 1.0, 0.0]   # test_gpu.jl:24


def bcc_small(seed=0, dims=(3, 3, 3)):
    na, t = M.lattice_model("bcc", pair_shells=3, pair_lmax=2, triangles=True, seed=seed)
    return na, t, dims


def bcc_tet_small(seed=0, dims=(3, 3, 3), isotropic=False):
    na, t = M.lattice_model("bcc", pair_shells=2, pair_lmax=2, triangles=True, tetrahedra=True,
                            isotropic=isotropic, seed=seed)
    return na, t, dims


def fixture_cases():
    """(name, n_cell_atoms, terms, dims, magmoms, field) — pair, triplet, general,
    Zeeman-only, inactive-site and lattice models (test_gpu.jl:181-194 analogues)."""
    return [
        ("dimer", 4, M.dimer_terms(-0.02), (1, 1, 1), None, None),
        ("chain", 1, M.chain_terms(0.05), (4, 1, 1), None, None),
        ("3body", 1, M.threebody_terms(0.05), (4, 2, 2), None, None),
        ("4body", 1, M.fourbody_terms(0.03), (4, 2, 2), None, None),
        ("dimer+field", 4, M.dimer_terms(-0.02), (2, 1, 1), ZMM_DIMER, ZB),
        ("zeeman-only", 2, [], (2, 2, 1), [1.0, 2.0], ZB),
        ("sc-onsite", *M.lattice_model("sc", pair_shells=2, pair_lmax=1, onsite_l2=True, seed=4), (4, 4, 4), None, None),
        ("bcc-l2-tri", *bcc_small(1), None, None),
        ("bcc-tet", *bcc_tet_small(2), None, None),
        ("fcc-l2", *M.lattice_model("fcc", pair_shells=2, pair_lmax=2, seed=5), (3, 3, 3), None, None),
    ]


def build_pair(na, terms, dims, mm=None, fl=None):
    H = S.TiledHamiltonian(na, terms, dims, magmoms=mm, field=fl)
    T = O.Tiled(na, terms, dims, magmoms=mm, field=fl)
    return H, T


def rand_config(n, seed):
    return S.random_config(n, np.random.default_rng(seed))
