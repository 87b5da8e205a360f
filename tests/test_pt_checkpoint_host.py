"""Host logic of ladder-run checkpoints (no device): the write cadence of
_ck_gpu_pt! (src/checkpoint.jl:428-437) — a write whenever `interval` sweeps
have passed since the last one, every write restarting the count, interval 0
meaning boundary writes only — and the run splits it forces."""
import pytest

from sce_amd.batch import _Checkpointer


def _cadence(interval, adapt, total):
    """The (sweeps done, wrote) chunk ends of a thermalisation phase driven as
    the ladder driver drives it."""
    ck = _Checkpointer("unused", interval)
    writes, ends, done = [], [], 0
    while done < total:
        n = ck.room(min(adapt - done % adapt, total - done))
        done += n
        ends.append(done)
        ck.tick(n, lambda path, d=done: writes.append(d))
    return ends, writes


def test_writes_land_on_interval_multiples_and_adaptation_boundaries_stay():
    ends, writes = _cadence(40, 50, 150)
    assert writes == [40, 80, 120]
    assert {50, 100, 150} <= set(ends)          # every adaptation boundary is a chunk end
    ends, writes = _cadence(10, 50, 150)
    assert writes == list(range(10, 151, 10))
    ends, writes = _cadence(120, 50, 150)
    assert writes == [120] and ends == [50, 100, 120, 150]


def test_interval_zero_writes_only_at_boundaries():
    ends, writes = _cadence(0, 50, 150)
    assert writes == [] and ends == [50, 100, 150]
    ck = _Checkpointer("unused", 0)
    got = []
    ck.write(lambda path: got.append(path))
    assert got == ["unused"]


def test_a_write_restarts_the_count():
    ck = _Checkpointer("p", 100)
    got = []
    ck.tick(60, lambda path: got.append(1))
    ck.write(lambda path: got.append(2))          # a boundary write
    ck.tick(60, lambda path: got.append(3))       # 60 since the boundary: no write
    assert got == [2]
    assert ck.room(1000) == 40


def test_negative_interval_rejected():
    with pytest.raises(ValueError):
        _Checkpointer("p", -1)
