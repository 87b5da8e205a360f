"""The allocation-free all-site gradient (sce_energy_gradient_noalloc) against
the allocating entry point it varies (sce_energy_gradient, the
gpu_energy_gradient! twin of src/gpu/gpu_gradient.jl:187-212): the same
kernels, so the same bits, on calls in a row through the C ABI — a larger
model, then a smaller one with inactive sites (whose rows must read zero
although the shared device array holds the previous call's values), then the
larger one again; sizes beyond the static array are refused before any device
work."""
import numpy as np
import pytest

import sce_amd as S
from sce_amd import _lib as L
from sce_amd import models as M
from sce_amd._lib import SceError
from fixtures import bcc_tet_small, rand_config

pytestmark = pytest.mark.gpu


def _noalloc(gH, cfg, lanes):
    f = L.lib().sce_energy_gradient_noalloc
    cfg = np.ascontiguousarray(cfg, np.float64)
    G = np.full((gH.host.n_sites, 3), np.nan)
    L.check(f(gH.handle, L.ptr(cfg), L.ptr(G), int(lanes)))
    return G


def _model(kind):
    if kind == "bcc6":
        return S.TiledHamiltonian(*bcc_tet_small(3, dims=(6, 6, 6)))
    if kind == "dimer":   # atoms 3-4 of the cell inactive
        return S.TiledHamiltonian(4, M.dimer_terms(-0.02), (2, 1, 1))
    raise KeyError(kind)


@pytest.mark.parametrize("lanes", [64, 128, 256])
def test_noalloc_gradient_equals_allocating_entry_point(lanes):
    Hs = {k: _model(k) for k in ("bcc6", "dimer")}
    gHs = {k: S.GPUTiledHamiltonian(H) for k, H in Hs.items()}
    for i, k in enumerate(("bcc6", "dimer", "bcc6")):
        H = Hs[k]
        cfg = rand_config(H.n_sites, 40 + i)
        ref = S.gpu_energy_gradient(gHs[k], cfg, lanes)
        got = _noalloc(gHs[k], cfg, lanes)
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), (k, i)
        if k == "dimer":
            assert not H.site_active.all() and (got[~H.site_active.astype(bool)] == 0.0).all()


def test_noalloc_rejects_sizes_beyond_the_static_array():
    """79 507 sites at lmax 6 need (3 + 49 + 3) doubles each, beyond the 4 Mi
    doubles of the array: a clear error, not a fault."""
    f = np.zeros(13)
    f[6] = 1.0
    terms = [M.MultipoleTerm(1e-3, [1], [(0, 0, 0)], [6], f)]
    H = S.TiledHamiltonian(1, terms, (43, 43, 43))
    gH = S.GPUTiledHamiltonian(H)
    with pytest.raises(SceError, match="static"):
        _noalloc(gH, rand_config(H.n_sites, 1), 128)
