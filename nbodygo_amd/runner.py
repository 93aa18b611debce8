"""Host side of the drop-in boundary in Python (ctypes): the reference's body collection, result
queue and computation runner (cmd/body/body_collection.go, cmd/runner/resultqueue.go,
cmd/runner/computation-runner.go) over libnbgpu.so, with the same phase order as the C++ driver
in nbodygo_amd/runner/:

    control polls (time scaling, R, deletes -> nb_remove_bodies, get/mod -> nb_get_body /
    nb_set_body) -> nb_step (force, detection, resolution, update: one cycle on the device)
    -> renderables into a result queue -> Cycle (nb_add_bodies + nb_cycle: compaction of removed
    bodies and the queued adds, on the device)

The device holds the body state for the whole run: it is loaded once, and the host keeps only
what the device does not (ids, and the radius / sun flag the renderables carry), compacted with
the Exists flags each render returns.  sync_to_host() pulls the whole state on demand.
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


class BodyCollection:
    """The host side of the sim's bodies (body_collection.go): a Bodies SoA mirror plus ids.

    Index i here is index i on the device.  The physical state in `b` is as of the last
    sync_to_host(); ids, radii and flags are kept current by the runner."""

    def __init__(self, bodies: Bodies, ids=None):
        self.b = bodies.copy()
        self.ids = (np.array([next_id() for _ in range(bodies.n)], np.int64) if ids is None
                    else np.asarray(ids, np.int64).copy())
        self._adds = []
        self.cycle_count = 0

    def count(self):
        return self.b.n

    def enqueue_add(self, body: Bodies):
        """Queues body for the next Cycle; its ids are drawn now, as NewBody draws them."""
        self._adds.append((body.copy(), np.array([next_id() for _ in range(body.n)], np.int64)))

    def has_adds(self):
        return bool(self._adds)

    def take_adds(self):
        """The queued adds as (Bodies, ids) in enqueue order; the queue is emptied."""
        adds, self._adds = self._adds, []
        return adds

    def cycle(self, R, keep=None, adds=(), fragments=0):
        """Mirrors nb_cycle: survivors (keep, default the Exists flags) in order, then the
        device's `fragments` fragment bodies (Elastic, radius 1, ids drawn here in creation
        order), then the adds ((Bodies, ids) in enqueue order, or Bodies, whose ids are drawn
        here), each group newest first with r = R (body_collection.go:253-296).  Returns True
        when the array changed."""
        from .gen import concat
        if keep is None:
            keep = (self.b.flags & L.NB_F_EXISTS).astype(bool)
        adds = [a if isinstance(a, tuple) else
                (a, np.array([next_id() for _ in range(a.n)], np.int64)) for a in adds]
        changed = not keep.all() or bool(adds) or fragments > 0
        if not keep.all():
            self.b = _select(self.b, keep)
            self.ids = self.ids[keep]
        if fragments:
            z = np.zeros(fragments)
            fr = Bodies(z, z, z, z, z, z, z, np.full(fragments, 1.0))  # fragcalc.go:81-82
            fr.rest[:] = R
            self.b = concat(self.b, fr)
            self.ids = np.concatenate(
                [self.ids, np.array([next_id() for _ in range(fragments)], np.int64)[::-1]])
        for a, ids in reversed(adds):
            a = _select(a, np.arange(a.n)[::-1])
            a.rest[:] = R
            self.b = concat(self.b, a)
            self.ids = np.concatenate([self.ids, ids[::-1]])
        self.cycle_count += 1
        return changed


def _select(b, mask):
    keys = ("x", "y", "z", "vx", "vy", "vz", "m", "r", "beh", "flags", "rest", "frag", "fstep")
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
        self._loaded = False
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
        """Pulls the whole device state into bc.b."""
        if self._stale and self._loaded and self.bc.count():
            s = self.sim.store()
            assert s.n == self.bc.count()
            self.bc.b = s
        self._stale = False

    def _index(self, id_):
        hit = np.nonzero(self.bc.ids == id_)[0]
        return hit

    def get_body(self, id_):
        """GetBody (body_collection.go:111-151): one body read from the device, or None."""
        self._push_if_dirty()
        hit = self._index(id_)
        if hit.size == 0:
            return None
        d = self.sim.get_body(int(hit[0]))
        return Bodies(*([d[k]] for k in ("x", "y", "z", "vx", "vy", "vz", "m", "r")),
                      beh=[d["beh"]], flags=[d["flags"]], rest=[d["rest"]], frag=[d["frag"]])

    def mod_body(self, id_, **fields):
        """ModBody (body_collection.go:154-191): the fields patched on the device."""
        self._push_if_dirty()
        hit = self._index(id_)
        for i in hit:
            self.sim.set_body(int(i), **fields)
            for k in ("r", "flags"):  # what the renderables read from the host side
                if k in fields:
                    getattr(self.bc.b, k)[i] = fields[k]
        if hit.size:
            self._stale = True
        return int(hit.size)

    def _process_deletes(self):
        """processDeletes (computation-runner.go:176-216) as nb_remove_bodies."""
        if self._del_box is None:
            return 0
        d, self._del_box = self._del_box, None
        self._push_if_dirty()
        if self.bc.count() == 0:
            return 0
        k = self.sim.remove_bodies(d)
        if k:
            self._stale = True
        self.stats["removed"] += k
        return k

    def _push_if_dirty(self):
        if self._dirty:
            self.sim.load(self.bc.b)
            self._loaded = True
            self.stats["loads"] += 1
        self._dirty = False

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
        adds = self.bc.take_adds()
        pending = self.sim.event_stats()["fragments_pending"] if self.sim.params.events else 0
        if gone or adds or pending:  # Cycle on the device, mirrored on the host side
            for a, _ in adds:
                self.sim.add_bodies(a)
            self.sim.cycle(self.R)
            self.bc.cycle(self.R, ex.astype(bool) if n else None, adds, fragments=pending)
            self._stale = True
            self.stats["compactions"] += 1
        else:
            self.bc.cycle_count += 1
        self.stats["computations"] += 1
        self.stats["wall_ms"] += (time.perf_counter() - t0) * 1e3
        return st
