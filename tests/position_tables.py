"""Shared helpers for the per-position table tests: the tables through the C
ABI (host builder, device build) and an independent numpy restatement."""
import ctypes as C

import numpy as np

from sce_amd import _lib as L


def host_tables(H):
    """(nbq, low or None, perm) from sce_position_tables_host (no device)."""
    sizes = np.zeros(3, np.int64)
    L.check(L.lib().sce_position_tables_host(C.byref(H.desc), None, None, None, L.ptr(sizes)))
    nbm, am, has = (int(x) for x in sizes)
    nbq = np.zeros(H.n_active * nbm, np.int32)
    low = np.zeros(max(H.n_active, 1), np.uint64)
    perm = np.zeros(H.n_active * am, np.uint16)
    L.check(L.lib().sce_position_tables_host(C.byref(H.desc), L.ptr(nbq), L.ptr(low), L.ptr(perm), L.ptr(sizes)))
    return nbq.reshape(H.n_active, nbm), (low[: H.n_active] if has else None), perm.reshape(H.n_active, am)


def device_tables(gH):
    H = gH.host
    sizes = np.zeros(3, np.int64)
    L.check(L.lib().sce_tables_position_tables(gH.handle, None, None, None, L.ptr(sizes)))
    nbm, am, has = (int(x) for x in sizes)
    nbq = np.zeros(H.n_active * nbm, np.int32)
    low = np.zeros(max(H.n_active, 1), np.uint64)
    perm = np.zeros(H.n_active * am, np.uint16)
    L.check(L.lib().sce_tables_position_tables(gH.handle, L.ptr(nbq), L.ptr(low), L.ptr(perm), L.ptr(sizes)))
    return nbq.reshape(H.n_active, nbm), (low[: H.n_active] if has else None), perm.reshape(H.n_active, am)


def numpy_tables(H, nbq):
    """The lower-colour masks and the permutation restated from the tables
    (colour classes, adjacency CSR, program ids), given the neighbour lists."""
    cs = H.color_sites - 1
    col_of = np.full(H.n_sites, -1)
    for c in range(H.n_colors):
        col_of[cs[H.color_ptr[c] - 1: H.color_ptr[c + 1] - 1]] = c
    low = np.zeros(H.n_active, np.uint64)
    for q in range(H.n_active):
        m = 0
        for j, n in enumerate(nbq[q]):
            if 0 <= col_of[n] < col_of[cs[q]]:
                m |= 1 << j
        low[q] = m
    am = max(int(np.diff(H.site_ptr).max()) if H.n_sites else 1, 1)
    perm = np.zeros((H.n_active, am), np.uint16)
    for q in range(H.n_active):
        s = cs[q]
        pid = H.site_prog[H.site_ptr[s] - 1: H.site_ptr[s + 1] - 1]
        order = np.argsort(pid, kind="stable")
        perm[q, order] = np.arange(len(pid))
    return low, perm
