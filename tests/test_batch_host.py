"""Host logic of run_concurrent: a batch may appear only once, every run's
result comes back in order, and an error raised in a worker thread reaches
the caller (dummy batches, no GPU)."""
import threading

import pytest

from sce_amd.batch import run_concurrent


class _Dummy:
    def __init__(self, name, fail=False):
        self.name, self.fail, self.thread = name, fail, None

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
