"""The placement of a chained sweep's proposals and deferred final (host only).

Sweeps chained within one call make the next sweep's proposals during the
current sweep and run the previous sweep's final reduction inside a walk
launch.  They are bit-identical to unchained sweeps exactly when every
proposal is made after its class's walk of the previous sweep and before its
class's walk of the sweep it is for, exactly once, keyed on the right sweep
index (the completed-sweep count plus an offset that counts a final still
pending at that launch), and no launch that advances the count carries one.
These tests check those conditions on the placements for random class sizes
(empty classes included) and every combination of the chaining flags."""
import numpy as np
import pytest

from sce_amd import gpu as G


def _cases():
    rng = np.random.default_rng(2026)
    yield [0, 5]                      # one class
    yield [0, 3, 3]                   # an empty class
    yield [0, 1, 9]                   # the smallest class first
    yield [0, 9, 10]                  # the smallest class last
    for _ in range(150):
        k = int(rng.integers(1, 18))
        sizes = rng.integers(0, 40, k)
        if rng.random() < 0.3:
            sizes[rng.integers(0, k)] = 0
        if sizes.sum() == 0:
            sizes[0] = 1
        yield [0] + list(np.cumsum(sizes))


@pytest.mark.parametrize("walk_props", [True, False])
@pytest.mark.parametrize("pending", [-1, 0, 1])
@pytest.mark.parametrize("next_prop", [True, False])
@pytest.mark.parametrize("carry_last", [True, False])
def test_proposal_placement(carry_last, next_prop, pending, walk_props):
    if carry_last and not walk_props:
        pytest.skip("not a state the sweep loops produce: a sweep carries its last class only when walks carry proposals")
    for cp in _cases():
        cp = np.asarray(cp)
        sizes = np.diff(cp)
        classes = [(int(cp[c]), int(sizes[c])) for c in range(len(sizes)) if sizes[c] > 0]
        K = len(classes)
        recs, kf = G.chain_plan(cp, next_prop, pending, carry_last, walk_props)
        if pending >= 0:
            want = min(range(K), key=lambda k: (classes[k][1], k))
            assert kf == want
        else:
            assert kf == -1
        lag = lambda j: 1 if kf >= 0 and j < kf else 0
        cover = {0: np.zeros(cp[-1], int), 1: np.zeros(cp[-1], int)}
        per_launch = {}
        for launch, lo, n, off in recs:
            assert n > 0
            nxt = off - 1 - lag(launch)            # 0: this sweep, 1: the next
            assert nxt in (0, 1), (cp, recs)
            cover[nxt][lo:lo + n] += 1
            per_launch[launch] = per_launch.get(launch, 0) + 1
            assert not (kf >= 0 and launch == kf)  # that launch advances the count
            first = [k for k in range(K) if classes[k][0] <= lo < classes[k][0] + classes[k][1]][0]
            last = [k for k in range(K) if classes[k][0] < lo + n <= classes[k][0] + classes[k][1]][0]
            if nxt == 0:   # this sweep's: before the class's own walk
                assert first == last == K - 1 and launch < K - 1
            else:          # the next sweep's: after the classes' walks
                assert launch > last
            if not walk_props:
                assert launch >= K
        want0 = np.zeros(cp[-1], int)
        if carry_last and K:
            lo, n = classes[-1]
            want0[lo:lo + n] = 1
        want1 = np.zeros(cp[-1], int)
        if next_prop:
            for k, (lo, n) in enumerate(classes):
                if not (walk_props and k == K - 1):
                    want1[lo:lo + n] = 1
        assert np.array_equal(cover[0], want0) and np.array_equal(cover[1], want1), (cp, recs)
        assert all(per_launch.get(j, 0) <= 2 for j in range(K)) and per_launch.get(K, 0) <= 1
