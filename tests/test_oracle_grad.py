"""Pins of the oracle's all-site energy gradient (the reference's
gpu_energy_gradient! / _gradient_lane_ref!, gpu/gpu_gradient.jl:92-279, and the
gradient row of gpu/grad_device.jl:30-66), CPU only: the gradient row against
central differences of the Z row and its tangency; the all-site gradient
against central differences of the total energy under tangent rotations of one
spin, the l=1 dimer closed form, zero rows at inactive sites, and lane-count
independence up to rounding."""
import math

import numpy as np
import pytest

import oracle as O
from sce_amd import models as M
from fixtures import build_pair, fixture_cases, rand_config

CASES = {c[0]: c for c in fixture_cases()}


def _tangents(u):
    a = np.array([1.0, 0.0, 0.0]) if abs(u[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    t1 = np.cross(u, a)
    t1 /= np.linalg.norm(t1)
    return t1, np.cross(u, t1)


@pytest.mark.parametrize("lmax", [0, 1, 2, 3, 4, 5, 6])
def test_gradient_row_is_the_tangent_derivative_of_the_row(lmax):
    rng = np.random.default_rng(lmax)
    h = 1e-6
    for _ in range(20):
        u = rng.standard_normal(3)
        u /= np.linalg.norm(u)
        g = O.grad_zlm_row(u, lmax)
        assert g.shape == ((lmax + 1) ** 2, 3)
        scale = max(1.0, float(np.abs(g).max()))
        assert np.abs(g @ u).max() <= 1e-13 * scale            # tangent projected
        for t in _tangents(u):
            zp = O.zlm_row((u + h * t) / np.linalg.norm(u + h * t), lmax)
            zm = O.zlm_row((u - h * t) / np.linalg.norm(u - h * t), lmax)
            np.testing.assert_allclose((zp - zm) / (2 * h), g @ t, atol=2e-8 * scale)
    if lmax == 0:
        assert np.all(O.grad_zlm_row(np.array([0.0, 0.6, 0.8]), 0) == 0.0)


def test_gradient_row_at_the_poles_and_axes_is_finite():
    for u in ([0, 0, 1.0], [0, 0, -1.0], [1.0, 0, 0], [0, -1.0, 0]):
        g = O.grad_zlm_row(np.array(u), 6)
        assert np.all(np.isfinite(g))
        assert np.abs(g @ np.array(u)).max() <= 1e-12


def _energy(T, cfg):
    return T.total_energy(T.zrows(cfg))


@pytest.mark.parametrize("name", ["dimer", "3body", "4body", "sc-onsite", "bcc-l2-tri", "bcc-tet", "fcc-l2",
                                  "dimer+field"])
def test_gradient_is_the_derivative_of_the_energy(name):
    _, na, terms, dims, mm, fl = CASES[name]
    _, T = build_pair(na, terms, dims, mm, fl)
    cfg = rand_config(T.n_sites, 17)
    G = T.energy_gradient(cfg, lanes=64)
    assert np.abs(np.sum(G * cfg, axis=1)).max() <= 1e-12 * max(1.0, np.abs(G).max())
    h = 1e-5
    rng = np.random.default_rng(2)
    for s in rng.choice(np.nonzero(T.site_active)[0], size=min(6, int(T.site_active.sum())), replace=False):
        for t in _tangents(cfg[s]):
            cp, cm = cfg.copy(), cfg.copy()
            cp[s] = (cfg[s] + h * t) / np.linalg.norm(cfg[s] + h * t)
            cm[s] = (cfg[s] - h * t) / np.linalg.norm(cfg[s] - h * t)
            fd = (_energy(T, cp) - _energy(T, cm)) / (2 * h)
            assert abs(fd - G[s] @ t) <= 1e-6 * max(1e-3, np.abs(G[s]).max()), (name, s, fd, G[s] @ t)


def test_dimer_gradient_closed_form_and_inactive_rows():
    J = -0.02
    _, T = build_pair(4, M.dimer_terms(J), (1, 1, 1))
    cfg = rand_config(T.n_sites, 5)
    G = T.energy_gradient(cfg, lanes=64)
    e1, e2 = cfg[0], cfg[1]
    np.testing.assert_allclose(G[0], J * (e2 - (e1 @ e2) * e1), atol=1e-15)
    np.testing.assert_allclose(G[1], J * (e1 - (e1 @ e2) * e2), atol=1e-15)
    inactive = np.nonzero(T.site_active == 0)[0]
    assert inactive.size and np.all(G[inactive] == 0.0)


def test_gradient_lane_count_changes_only_rounding():
    _, na, terms, dims, mm, fl = CASES["bcc-tet"]
    _, T = build_pair(na, terms, dims, mm, fl)
    cfg = rand_config(T.n_sites, 8)
    z = T.zrows(cfg)
    G1 = T.energy_gradient(cfg, z, lanes=1)
    for ws in (64, 128, 256):
        np.testing.assert_allclose(T.energy_gradient(cfg, z, lanes=ws), G1, rtol=1e-12, atol=1e-16)
