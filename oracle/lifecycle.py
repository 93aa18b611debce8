"""CPU restatement of the reference's collection bookkeeping -- TEST INFRASTRUCTURE ONLY.

Checks the device lifecycle calls (nb_remove_bodies, nb_add_bodies + nb_cycle, nb_get_body /
nb_set_body).  Pure Python / numpy over a Bodies-like SoA (attributes x..vz, m, r, beh, flags,
rest, frag).

remove_indices   ComputationRunner.processDeletes, cmd/runner/computation-runner.go:176-216.
                 Pinned by hand-worked walks of the Go code (tests/test_lifecycle.py), by brute
                 force against a literal transcription of the loop with its closure, and by the
                 invariants the walk fixes (one removal per merged stride run).
cycle            BodyCollection.Cycle, cmd/body/body_collection.go:253-296: survivors keep their
                 order, then the queued adds newest first (events are pushed at the front of the
                 list, :100, and Cycle walks it from the front, :273-279 / :286-290), each add's R
                 set to the cycle's R.  Pinned by hand-worked cases.
"""
import numpy as np

EXISTS, FRAGMENTING, SUN, PINNED = 1, 2, 4, 8
FIELDS = ("x", "y", "z", "vx", "vy", "vz", "m", "r", "beh", "flags", "rest", "frag")


def remove_indices(exists, pinned, del_cnt):
    """The body indices processDeletes sets to not exist (computation-runner.go:176-216).

    -1 removes every existing body (pinned ones too, :183-188).  Otherwise step = 1 when more
    are asked than the collection holds, else Count() // del_cnt (:191-195); the walk raises
    shouldRemove at every multiple of step (:198-200) and clears it at the first existing,
    unpinned body at or after that point (:202-205).  Reaching del_cnt returns from one callback
    only (:207-209), so the walk goes on: with the array's tail the walk can remove more than
    asked.
    """
    n = len(exists)
    if del_cnt == -1:
        return [i for i in range(n) if exists[i]]
    if del_cnt <= 0:  # 0 divides by zero there (:195) and < -1 is not a request; no removal
        return []
    step = 1 if del_cnt > n else n // del_cnt
    out = []
    should = False
    for it in range(n):
        if it % step == 0:
            should = True
        if should and not pinned[it] and exists[it]:
            out.append(it)
            should = False
    return out


def apply_removals(b, idx):
    """SetNotExists (body.go:93-96) on each index: Exists cleared, mass 0."""
    for i in idx:
        b.flags[i] &= ~np.uint8(EXISTS)
        b.m[i] = 0.0


def cycle(b, adds, R):
    """(survivors in order) + (the bodies of adds, a list of SoAs in enqueue order, newest
    first, rest = R) as a dict of new arrays."""
    keep = (b.flags & EXISTS) != 0
    out = {k: getattr(b, k)[keep].copy() for k in FIELDS}
    for a in reversed(adds):
        for k in FIELDS:
            v = getattr(a, k)[::-1].copy()
            if k == "rest":
                v[:] = R
            out[k] = np.concatenate([out[k], v])
    return out
