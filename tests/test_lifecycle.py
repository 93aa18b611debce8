"""CPU checks of the collection-lifecycle oracle (oracle/lifecycle.py) and of the host-side
mirrors the drivers keep.

processDeletes (computation-runner.go:176-216): hand-worked walks of the Go loop, a literal
closure-style transcription on random small inputs, and the invariants the walk fixes.
Cycle (body_collection.go:253-296): a hand-worked compaction with adds newest first.
"""
import numpy as np
import pytest

from oracle import lifecycle as LC
from nbodygo_amd import gen
from nbodygo_amd import runner as R


def test_delete_walk_hand_worked():
    ex = np.ones(10, bool)
    pin = np.zeros(10, bool)
    # step = 10 // 3 = 3: stride points 0, 3, 6, 9; the fourth removal is the closure quirk
    assert LC.remove_indices(ex, pin, 3) == [0, 3, 6, 9]
    pin[3] = True  # a pinned body passes the removal on to the next live body
    assert LC.remove_indices(ex, pin, 3) == [0, 4, 6, 9]
    ex[0] = False
    assert LC.remove_indices(ex, pin, 3) == [1, 4, 6, 9]
    assert LC.remove_indices(ex, pin, 50) == [1, 2, 4, 5, 6, 7, 8, 9]  # more than exist: step 1
    assert LC.remove_indices(ex, pin, -1) == [1, 2, 3, 4, 5, 6, 7, 8, 9]  # pinned ones too
    assert LC.remove_indices(ex, pin, 0) == []
    # n = 7, 2 asked: step 3, stride points 0, 3, 6 -> three removed
    assert LC.remove_indices(np.ones(7, bool), np.zeros(7, bool), 2) == [0, 3, 6]
    # a stride run with no live body merges into the next: n = 9, step 3, bodies 0-3 gone
    ex = np.array([0, 0, 0, 0, 1, 1, 1, 1, 1], bool)
    assert LC.remove_indices(ex, np.zeros(9, bool), 3) == [4, 6]
    # an empty collection: 5 > 0 gives step 1 and nothing to walk
    assert LC.remove_indices(np.zeros(0, bool), np.zeros(0, bool), 5) == []


def _go_closure_walk(exists, pinned, del_cnt):
    """The loop as the Go source states it: a callback per body with the state in the enclosing
    scope, SetNotExists on the callback's own copy, and `return` leaving only the callback."""
    ex0 = list(exists)
    exists = list(exists)
    state = {"iter": 0, "should": False, "removed": 0}
    n = len(exists)
    step = 1 if del_cnt > n else n // del_cnt

    def consumer(i):
        if state["iter"] % step == 0:
            state["should"] = True
        state["iter"] += 1
        if state["should"] and not pinned[i] and exists[i]:
            exists[i] = False
            state["should"] = False
            state["removed"] += 1
            if state["removed"] >= del_cnt:
                return

    for i in range(n):  # IterateOnce
        consumer(i)
    return [i for i in range(n) if ex0[i] and not exists[i]]


def test_delete_walk_matches_closure_transcription():
    rng = np.random.default_rng(0)
    for _ in range(400):
        n = int(rng.integers(0, 60))
        ex = rng.random(n) < rng.uniform(0.3, 1.0)
        pin = rng.random(n) < rng.uniform(0.0, 0.4)
        d = int(rng.integers(1, n + 5))
        assert LC.remove_indices(ex, pin, d) == _go_closure_walk(ex, pin, d), (n, d)


@pytest.mark.parametrize("n,d", [(10, 3), (1000, 7), (1000, 1000), (999, 1000), (12345, 100)])
def test_delete_walk_invariants(n, d):
    ex = np.ones(n, bool)
    pin = np.zeros(n, bool)
    idx = LC.remove_indices(ex, pin, d)
    step = 1 if d > n else n // d
    # every stride point removes exactly its own body when all exist and none is pinned
    assert idx == list(range(0, n, step))
    assert len(idx) >= min(d, n)


def test_cycle_hand_worked():
    b = gen.uniform_sphere(4, seed=1)
    b.x[:] = [0, 1, 2, 3]
    b.flags[1] = 0
    a1 = gen.uniform_sphere(2, seed=2)
    a1.x[:] = [10, 11]
    a2 = gen.uniform_sphere(1, seed=3)
    a2.x[:] = [20]
    out = LC.cycle(b, [a1, a2], 0.25)
    assert list(out["x"]) == [0, 2, 3, 20, 11, 10]
    assert list(out["rest"][3:]) == [0.25] * 3
    assert list(out["rest"][:3]) == list(b.rest[[0, 2, 3]])


def test_host_mirror_cycle_matches_oracle():
    b = gen.uniform_sphere(6, seed=1)
    bc = R.BodyCollection(b)
    ids = bc.ids.copy()
    keep = np.ones(6, bool)
    keep[[1, 4]] = False
    add = gen.uniform_sphere(2, seed=2)
    add2 = gen.uniform_sphere(3, seed=4)
    b2 = b.copy()
    b2.flags[~keep] = 0
    exp = LC.cycle(b2, [add, add2], 0.5)
    assert bc.cycle(0.5, keep, [add, add2])
    assert bc.count() == 9 and list(bc.ids[:4]) == list(ids[[0, 2, 3, 5]])
    for k in ("x", "vy", "m", "r", "rest"):
        np.testing.assert_array_equal(getattr(bc.b, k), exp[k])
    assert not bc.cycle(0.5, np.ones(9, bool), [])
