"""Host side of the drop-in boundary in Python (ctypes): the reference's body collection, result
queue and computation runner (cmd/body/body_collection.go, cmd/runner/resultqueue.go,
cmd/runner/computation-runner.go) over libnbgpu.so, with the same phase order as the C++ driver
in nbodygo_amd/runner/:

    control polls (time scaling, R, deletes, get/mod requests)
    -> nb_step (force, detection, resolution, update: one cycle on the device)
    -> renderables into a result queue -> Cycle (compaction of removed bodies, adds)

The device holds the body state between cycles; the host copy is pulled (nb_store) only when a
host-side decision needs it and pushed (nb_load) after the host changed it.
"""
import collections
import dataclasses
import itertools
import time

import numpy as np

from . import _lib as L
from .sim import Bodies, NBodyGPU

_ids = itertools.count()


def next_id():
    """body.NextId (body.go:316-329)."""
    return next(_ids)


@dataclasses.dataclass
class Renderable:
    """renderable.go:9-40."""
    id: int
    exists: bool
    x: float = 0.0
    y: float = 0.0
    z: float = 0.0
    radius: float = 0.0
    is_sun: bool = False


@dataclasses.dataclass
class ResultQueue:
    queue_num: int
    queue: list = dataclasses.field(default_factory=list)

    def add(self, r):
        self.queue.append(r)


class ResultQueueHolder:
    """A bounded FIFO of per-cycle result queues (resultqueue.go)."""

    def __init__(self, max_queues):
        self.max_queues = max_queues
        self._q = collections.deque()
        self._num = 0

    def new_result_queue(self):
        if len(self._q) >= self.max_queues:
            return None
        q = ResultQueue(self._num)
        self._num += 1
        return q

    def add(self, q):
        self._q.append(q)

    def next(self):
        return self._q.popleft() if self._q else None

    def resize(self, max_queues):
        if max_queues == self.max_queues:
            return False
        self.max_queues = max_queues
        return True


def remove_indices(exists, pinned, del_cnt):
    """The bodies processDeletes removes (computation-runner.go:176-216), including the closure
    quirk: reaching del_cnt returns from one callback only, so the walk goes on and the next
    stride point removes again."""
    n = len(exists)
    out = []
    if del_cnt == -1:
        return [i for i in range(n) if exists[i]]
    if del_cnt <= 0:
        return out
    step = 1 if del_cnt > n else n // del_cnt
    should = False
    for it in range(n):
        if it % step == 0:
            should = True
        if should and not pinned[it] and exists[it]:
            out.append(it)
            should = False
    return out


class BodyCollection:
    """The host copy of the sim's bodies (body_collection.go) as a Bodies SoA plus ids."""

    def __init__(self, bodies: Bodies, ids=None):
        self.b = bodies.copy()
        self.ids = (np.array([next_id() for _ in range(bodies.n)], np.int64) if ids is None
                    else np.asarray(ids, np.int64).copy())
        self._adds = []
        self.cycle_count = 0

    def count(self):
        return self.b.n

    def enqueue_add(self, body: Bodies):
        self._adds.append(body.copy())

    def has_adds(self):
        return bool(self._adds)

    def cycle(self, R):
        """Compacts away removed bodies and appends queued adds (each gets r = R)
        (body_collection.go:253-296).  Returns True when the array changed."""
        from .gen import concat
        keep = (self.b.flags & L.NB_F_EXISTS).astype(bool)
        changed = not keep.all() or bool(self._adds)
        if not keep.all():
            self.b = _select(self.b, keep)
            self.ids = self.ids[keep]
        for a in self._adds:
            a.rest[:] = R
            self.b = concat(self.b, a)
            self.ids = np.concatenate([self.ids, [next_id() for _ in range(a.n)]])
        self._adds = []
        self.cycle_count += 1
        return changed


def _select(b, mask):
    keys = ("x", "y", "z", "vx", "vy", "vz", "m", "r", "beh", "flags", "rest", "frag")
    return Bodies(*(getattr(b, k)[mask].copy() for k in keys))


class ComputationRunner:
    """computation-runner.go:24-326 over the device.  worker_cnt is kept for the reference's
    signature; the device replaces the worker pool, and the direct sum always runs."""

    def __init__(self, worker_cnt, time_scaling, result_queue_holder, bc: BodyCollection,
                 **params):
        self.worker_cnt = worker_cnt
        self.time_scaling = time_scaling
        self.R = 1.0
        self.rqh = result_queue_holder
        self.bc = bc
        self.sim = NBodyGPU(**params)
        self._time_scale_box = None
        self._R_box = None
        self._del_box = None
        self._dirty = True
        self._stale = False
        self.stats = collections.Counter()

    def close(self):
        self.sim.close()

    # control surface (each a one-slot mailbox polled at the start of a cycle)
    def set_time_scaling(self, t):
        self._time_scale_box = float(t)

    def set_coefficient_of_restitution(self, R):
        self._R_box = float(R)

    def remove_bodies(self, n):
        self._del_box = int(n)

    def sync_to_host(self):
        if self._stale:
            s = self.sim.store()
            for k in ("x", "y", "z", "vx", "vy", "vz", "flags", "rest"):
                getattr(self.bc.b, k)[:] = getattr(s, k)
            gone = ~(self.bc.b.flags & L.NB_F_EXISTS).astype(bool)
            self.bc.b.m[gone] = 0.0
            self._stale = False

    def get_body(self, id_):
        self.sync_to_host()
        hit = np.nonzero(self.bc.ids == id_)[0]
        if hit.size == 0:
            return None
        return _select(self.bc.b, np.isin(np.arange(self.bc.count()), hit[:1]))

    def mod_body(self, id_, **fields):
        self.sync_to_host()
        hit = np.nonzero(self.bc.ids == id_)[0]
        for k, v in fields.items():
            getattr(self.bc.b, k)[hit] = v
        if hit.size:
            self._dirty = True
        return int(hit.size)

    def _process_deletes(self):
        if self._del_box is None:
            return 0
        d, self._del_box = self._del_box, None
        self.sync_to_host()
        b = self.bc.b
        idx = remove_indices((b.flags & L.NB_F_EXISTS).astype(bool),
                             (b.flags & L.NB_F_PINNED).astype(bool), d)
        for i in idx:  # SetNotExists (body.go:93-96)
            b.flags[i] &= ~np.uint8(L.NB_F_EXISTS)
            b.m[i] = 0.0
        if idx:
            self._dirty = True
        return len(idx)

    def _push_if_dirty(self):
        if self._dirty and self.bc.count() > 0:
            self.sim.load(self.bc.b)
            self.stats["loads"] += 1
        self._dirty = False
        self._stale = False

    def run_one_computation(self):
        """One cycle in the reference's phase order (computation-runner.go:267-326)."""
        t0 = time.perf_counter()
        self.stats["iterations"] += 1
        if self._time_scale_box is not None:
            self.time_scaling, self._time_scale_box = self._time_scale_box, None
        if self._R_box is not None:
            self.R, self._R_box = self._R_box, None
        self._process_deletes()
        rq = self.rqh.new_result_queue()
        if rq is None:
            self.stats["skipped_no_queue"] += 1
            return None
        self._push_if_dirty()
        n = self.bc.count()
        st = None
        if n:
            st = self.sim.step(self.time_scaling, self.R, 1, allow_overflow=True)
            self._stale = True
            xyz, ex = self.sim.render()
            b = self.bc.b
            for i in range(n):
                rq.add(Renderable(int(self.bc.ids[i]), bool(ex[i]), *map(float, xyz[i]),
                                  float(b.r[i]), bool(b.flags[i] & L.NB_F_SUN)))
            gone = not ex.all()
        else:
            gone = False
        self.rqh.add(rq)
        if gone or self.bc.has_adds():
            self.sync_to_host()
        if self.bc.cycle(self.R):
            self._dirty = True
        self.stats["computations"] += 1
        self.stats["wall_ms"] += (time.perf_counter() - t0) * 1e3
        return st
