"""The host builder of the per-position tables (sce_position_tables_host) against
an independent restatement: every neighbour slot is a co-instance site of the
position's site (or the site itself as padding) and every co-instance site
appears; the lower-colour masks and the raw -> pid-sorted permutation equal
the numpy restatement from the colour classes and adjacency programs."""
import numpy as np
import pytest

import sce_amd as S
from fixtures import fixture_cases
from position_tables import host_tables, numpy_tables


@pytest.mark.parametrize("case", fixture_cases(), ids=lambda c: c[0])
def test_host_position_tables(case):
    name, na, terms, dims, mm, fl = case
    H = S.TiledHamiltonian(na, terms, dims, magmoms=mm, field=fl)
    nbq, low, perm = host_tables(H)
    cs = H.color_sites - 1
    for q in range(H.n_active):
        s = cs[q]
        insts = H.site_inst[H.site_ptr[s] - 1: H.site_ptr[s + 1] - 1]
        co = set()
        for i in insts:
            co.update(int(x) - 1 for x in H.inst_sites[H.inst_ptr[i - 1] - 1: H.inst_ptr[i] - 1])
        co.discard(s)
        assert co <= set(int(x) for x in nbq[q])
        assert set(int(x) for x in nbq[q]) <= co | {s}
    low_ref, perm_ref = numpy_tables(H, nbq)
    if low is not None:
        assert np.array_equal(low, low_ref)
    assert np.array_equal(perm[:, : perm_ref.shape[1]], perm_ref[:, : perm.shape[1]])
