"""Host logic of run_concurrent: a batch or chain may appear only once, every
run's result comes back in order, functions run with their arguments, and an
error raised in a worker thread reaches the caller (dummy batches, no GPU)."""
import threading

import pytest

from sce_amd.batch import run_concurrent


class _Dummy:
    def __init__(self, name, fail=False):
        self.name, self.fail, self.thread = name, fail, None
        self._graph, self._persistent, self.graph_calls = True, False, []

    def set_graph(self, on):
        self.graph_calls.append(on)
        self._graph = on

    def run(self, n, lanes, **kw):
        self.thread = threading.get_ident()
        if self.fail:
            raise RuntimeError(f"{self.name} failed")
        return (self.name, n, lanes, kw)


def test_results_in_order_on_their_own_threads():
    a, b = _Dummy("a"), _Dummy("b")
    out = run_concurrent([(a, (3, 16), {"renorm_interval": 2}), (b, (4, 64), {})])
    assert out == [("a", 3, 16, {"renorm_interval": 2}), ("b", 4, 64, {})]
    assert a.thread != threading.get_ident() and b.thread != threading.get_ident()


def test_a_batch_appears_once():
    a = _Dummy("a")
    with pytest.raises(ValueError):
        run_concurrent([(a, (1, 16), {}), (a, (1, 16), {})])
    assert a.thread is None


def test_worker_errors_reach_the_caller():
    with pytest.raises(RuntimeError, match="b failed"):
        run_concurrent([(_Dummy("a"), (1, 16), {}), (_Dummy("b", fail=True), (1, 16), {})])


def test_functions_and_their_chains():
    seen = []

    def sweep(chain, n, lanes=64):
        seen.append((chain, n, lanes, threading.get_ident()))
        return chain * n

    out = run_concurrent([(sweep, (2, 5), {}), (sweep, (3, 7), {"lanes": 16})])
    assert out == [10, 21]
    assert sorted(x[:3] for x in seen) == [(2, 5, 64), (3, 7, 16)]
    assert all(x[3] != threading.get_ident() for x in seen)
    chain = object()
    with pytest.raises(ValueError):
        run_concurrent([(sweep, (chain, 1), {}), (sweep, (chain, 2), {})])


def test_graph_replay_off_during_the_run_and_restored():
    a, b = _Dummy("a"), _Dummy("b")
    b._graph = False
    seen = {}

    class _Probe(_Dummy):
        def run(self, n, lanes, **kw):
            seen["graph_during"] = (a._graph, b._graph)
            return super().run(n, lanes, **kw)
    p = _Probe("p")
    run_concurrent([(a, (1, 16), {}), (b, (1, 16), {}), (p, (1, 16), {})])
    assert seen["graph_during"] == (False, False)
    assert a._graph and not b._graph and p._graph
    assert a.graph_calls == [False, True] and b.graph_calls == []
    f = _Dummy("f", fail=True)
    with pytest.raises(RuntimeError):
        run_concurrent([(f, (1, 16), {})])
    assert f._graph and f.graph_calls == [False, True]


def test_persistent_targets_are_refused():
    a, b = _Dummy("a"), _Dummy("b")
    b._persistent = True
    with pytest.raises(ValueError, match="persistent"):
        run_concurrent([(a, (1, 16), {}), (b, (1, 16), {})])
    assert a.thread is None and a.graph_calls == []
