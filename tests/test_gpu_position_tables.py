"""Device-side table build: the per-position tables sce_tables_create makes on
the device (neighbour sites, lower-colour masks, raw -> pid-sorted adjacency
permutation) equal the host builder bit for bit on every fixture model, the
Nd2Fe14B setup cell and the benchmark's bcc 32^3; a table that is not a
translation-invariant tiling is rejected by the device verification."""
import ctypes as C
import os

import numpy as np
import pytest

import sce_amd as S
from sce_amd import _lib as L
from sce_amd import models as M
from fixtures import fixture_cases
from position_tables import device_tables, host_tables

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _cases():
    out = [(c[0], S.TiledHamiltonian(c[1], c[2], c[3], magmoms=c[4], field=c[5])) for c in fixture_cases()]
    setup = S.read_setup(os.path.join(HERE, "golden", "nd2fe14b.toml"))
    na, terms = S.setup_model(setup, nbody=3, cutoff=3.5, lmax=[2] * 8 + [0], seed=3)
    out.append(("nd2fe14b-nbody3", S.TiledHamiltonian(na, terms, (2, 2, 2), magmoms=S.setup_magmoms(setup),
                                                       field=(0.0, 0.0, 1.0))))
    na, terms, dims = M.config_model("bcc32")
    out.append(("bcc32", S.TiledHamiltonian(na, terms, dims)))
    return out


@pytest.mark.parametrize("name, H", _cases(), ids=lambda v: v if isinstance(v, str) else "")
def test_device_position_tables_equal_host(name, H):
    gH = S.GPUTiledHamiltonian(H)
    hn, hl, hp = host_tables(H)
    dn, dl, dp = device_tables(gH)
    assert np.array_equal(dn, hn)
    assert (hl is None) == (dl is None)
    if hl is not None:
        assert np.array_equal(dl, hl)
    assert np.array_equal(dp, hp)


def test_device_verification_rejects_non_tiling_tables():
    H = S.TiledHamiltonian(1, M.threebody_terms(0.05), (4, 2, 2))
    d = L.TablesDesc.from_buffer_copy(H.desc)
    col = np.array(H.site_col, copy=True)
    j = int(np.nonzero(col < 0)[0][5])
    col[j] = -((-col[j]) % H.n_sites + 1)  # point one triplet at a wrong neighbour
    d.site_col = col.ctypes.data_as(L.i32p)
    h = C.c_void_p()
    rc = L.lib().sce_tables_create(C.byref(d), C.byref(h))
    assert rc == -6 and "translated stencil" in L.last_error()
