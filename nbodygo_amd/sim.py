"""Python driver of the C-ABI: body arrays and a context wrapper.

`Bodies` is the SoA form of the reference's body array (Body fields X..Mass, CollisionBehavior,
Exists/fragmenting/IsSun/Pinned, Body.r and FragFactor; cmd/body/body.go:34-52).  `NBodyGPU` owns
one nb_ctx and exposes the entry points of include/nbody_gpu.h with numpy arrays.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from ._lib import check, lib


@dataclass
class Bodies:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    vz: np.ndarray
    m: np.ndarray
    r: np.ndarray
    beh: np.ndarray = None
    flags: np.ndarray = None
    rest: np.ndarray = None
    frag: np.ndarray = None
    fstep: np.ndarray = None  # Body.FragStep (fragcalc.go:72)

    def __post_init__(self):
        for k in ("x", "y", "z", "vx", "vy", "vz", "m", "r"):
            setattr(self, k, np.ascontiguousarray(getattr(self, k), dtype=np.float64))
        n = self.n
        if self.beh is None:
            self.beh = np.full(n, L.NB_CB_ELASTIC, np.int32)
        if self.flags is None:
            self.flags = np.full(n, L.NB_F_EXISTS, np.uint8)
        self.beh = np.ascontiguousarray(self.beh, dtype=np.int32)
        self.flags = np.ascontiguousarray(self.flags, dtype=np.uint8)
        if self.rest is None:
            self.rest = np.ones(n, np.float64)
        self.rest = np.ascontiguousarray(self.rest, dtype=np.float64)
        if self.frag is None:
            self.frag = np.zeros(n, np.float64)
        self.frag = np.ascontiguousarray(self.frag, dtype=np.float64)
        if self.fstep is None:
            self.fstep = np.zeros(n, np.float64)
        self.fstep = np.ascontiguousarray(self.fstep, dtype=np.float64)
        for k in ("y", "z", "vx", "vy", "vz", "m", "r", "beh", "flags", "rest", "frag", "fstep"):
            if getattr(self, k).shape != (n,):
                raise ValueError("Bodies.%s has shape %s, expected (%d,)" %
                                 (k, getattr(self, k).shape, n))

    @property
    def n(self):
        return int(self.x.shape[0])

    def copy(self):
        return Bodies(*(getattr(self, k).copy() for k in
                        ("x", "y", "z", "vx", "vy", "vz", "m", "r", "beh", "flags", "rest",
                         "frag", "fstep")))

    def pos(self):
        return np.stack([self.x, self.y, self.z], axis=1)

    def vel(self):
        return np.stack([self.vx, self.vy, self.vz], axis=1)

    def active(self):
        return ((self.flags & L.NB_F_EXISTS) != 0) & ((self.flags & L.NB_F_FRAGMENTING) == 0)


def _dp(a):
    return a.ctypes.data_as(L._dp) if a is not None else None


PRECISIONS = {"fp32": L.NB_PREC_FP32, "fp64": L.NB_PREC_FP64}


class NBodyGPU:
    """One device context (nb_ctx)."""

    def __init__(self, G=6.673e-11, eps=0.0, overlap_cutoff=True, collisions=True,
                 precision="fp32", device=0, pair_capacity=0, ipl=0, tile_j=0, jsplit=0,
                 block=0, variant=0, events=False, resolve=0):
        p = L.NbParams()
        check(lib.nb_params_init(C.byref(p)))
        p.G = float(G)
        p.eps = float(eps)
        p.overlap_cutoff = int(bool(overlap_cutoff))
        p.collisions = int(bool(collisions))
        p.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        p.device = int(device)
        p.pair_capacity = int(pair_capacity)
        p.ipl, p.tile_j, p.jsplit, p.block, p.variant = ipl, tile_j, jsplit, block, variant
        p.events = int(bool(events))
        p.resolve = int(resolve)
        self.params = p
        self._ctx = C.c_void_p()
        check(lib.nb_create(C.byref(p), C.byref(self._ctx)), None)
        self.n = 0

    def close(self):
        if self._ctx:
            lib.nb_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        return check(rc, self._ctx)

    def load(self, b: Bodies):
        self._check(lib.nb_load(self._ctx, b.n, _dp(b.x), _dp(b.y), _dp(b.z), _dp(b.vx),
                                _dp(b.vy), _dp(b.vz), _dp(b.m), _dp(b.r),
                                b.beh.ctypes.data_as(L._i32p), b.flags.ctypes.data_as(L._u8p),
                                _dp(b.rest), _dp(b.frag)))
        self.n = b.n
        if self.params.events:
            self._check(lib.nb_set_frag_steps(self._ctx, _dp(b.fstep)))
        return self

    def step(self, dt, R=1.0, nsteps=1, allow_overflow=False):
        st = L.NbStats()
        rc = lib.nb_step(self._ctx, float(dt), float(R), int(nsteps), C.byref(st))
        if not (allow_overflow and rc == L.NB_E_PAIR_OVERFLOW):
            self._check(rc)
        return st.as_dict()

    def forces(self):
        f = np.empty((3, self.n), np.float64)
        self._check(lib.nb_compute_forces(self._ctx, _dp(f[0]), _dp(f[1]), _dp(f[2])))
        return f.T.copy()

    def accel_fp64(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        a = np.empty((3, idx.shape[0]), np.float64)
        self._check(lib.nb_accel_fp64(self._ctx, idx.ctypes.data_as(L._i64p), idx.shape[0],
                                      _dp(a[0]), _dp(a[1]), _dp(a[2])))
        return a.T.copy()

    def pairs(self):
        cnt = C.c_int64()
        self._check(lib.nb_get_pairs(self._ctx, None, 0, C.byref(cnt)))
        k = cnt.value
        out = np.empty((max(k, 1), 2), np.int32)
        self._check(lib.nb_get_pairs(self._ctx, out.ctypes.data_as(L._i32p), k, C.byref(cnt)))
        return out[:min(k, cnt.value)].copy()

    def store(self):
        """The whole device state as Bodies (nb_store_bodies)."""
        n = self.n
        f = [np.empty(n, np.float64) for _ in range(8)]
        beh = np.empty(n, np.int32)
        flags = np.empty(n, np.uint8)
        rest = np.empty(n, np.float64)
        frag = np.empty(n, np.float64)
        fstep = np.empty(n, np.float64)
        self._check(lib.nb_store_bodies(self._ctx, *(_dp(a) for a in f),
                                        beh.ctypes.data_as(L._i32p),
                                        flags.ctypes.data_as(L._u8p), _dp(rest), _dp(frag),
                                        _dp(fstep)))
        return Bodies(*f, beh=beh, flags=flags, rest=rest, frag=frag, fstep=fstep)

    # -- Barnes-Hut octree (cmd/body/octree.go with Build wired in) ---------------------------
    def bh_enable(self, theta=0.3, softening=1e-4, max_depth=21, group=0, mode="exact", term=0):
        """The reference's octree force (FP64 contexts).  mode "exact": per-body walks with the
        reference's test and FP64 terms in its order; "fast": one walk per 64-body group that
        opens what any of them would, FP32 terms.  term (exact mode): 0 by occupancy, 1 three
        divisions, 2 one shared reciprocal (bit-identical to 1)."""
        p = L.NbBhParams()
        check(lib.nb_bh_params_init(C.byref(p)))
        p.theta, p.softening, p.max_depth = float(theta), float(softening), int(max_depth)
        p.group = int(group)
        p.mode = {"exact": 0, "fast": 1}[mode] if isinstance(mode, str) else int(mode)
        p.term = int(term)
        self._check(lib.nb_bh_enable(self._ctx, C.byref(p)))

    def bh_disable(self):
        self._check(lib.nb_bh_disable(self._ctx))

    def bh_stats(self):
        st = L.NbBhStats()
        self._check(lib.nb_bh_get_stats(self._ctx, C.byref(st)))
        return st.as_dict()

    def bh_depths(self):
        d = np.empty(max(self.n, 1), np.int32)
        self._check(lib.nb_bh_get_depths(self._ctx, d.ctypes.data_as(L._i32p)))
        return d[:self.n].copy()

    # -- collection lifecycle on the device (body_collection.go, computation-runner.go) -------
    def add_bodies(self, b: Bodies):
        """Queues b for the next cycle() (BodyCollection.Enqueue)."""
        self._check(lib.nb_add_bodies(self._ctx, b.n, _dp(b.x), _dp(b.y), _dp(b.z), _dp(b.vx),
                                      _dp(b.vy), _dp(b.vz), _dp(b.m), _dp(b.r),
                                      b.beh.ctypes.data_as(L._i32p),
                                      b.flags.ctypes.data_as(L._u8p), _dp(b.frag),
                                      _dp(b.fstep)))

    def cycle(self, R=1.0):
        """Compaction of removed bodies, then the queued adds with restitution R
        (BodyCollection.Cycle).  Returns the new body count."""
        n = C.c_int64()
        self._check(lib.nb_cycle(self._ctx, float(R), C.byref(n)))
        self.n = n.value
        return self.n

    def remove_bodies(self, count):
        """processDeletes' removal walk (count, or -1 for all).  Returns the number removed."""
        k = C.c_int64()
        self._check(lib.nb_remove_bodies(self._ctx, int(count), C.byref(k)))
        return k.value

    def event_stats(self):
        st = L.NbEventStats()
        self._check(lib.nb_get_event_stats(self._ctx, C.byref(st)))
        return st.as_dict()

    def get_body(self, i):
        out = L.NbBody()
        self._check(lib.nb_get_body(self._ctx, int(i), C.byref(out)))
        return out.as_dict()

    def set_body(self, i, **fields):
        """Patches body i (ModBody): the fields given, the rest kept."""
        cur = L.NbBody()
        self._check(lib.nb_get_body(self._ctx, int(i), C.byref(cur)))
        for k, v in fields.items():
            if k not in cur.as_dict():
                raise KeyError(k)
            setattr(cur, k, v)
        self._check(lib.nb_set_body(self._ctx, int(i), C.byref(cur)))

    def render(self):
        xyz = np.empty((self.n, 3), np.float32)
        ex = np.empty(self.n, np.uint8)
        self._check(lib.nb_store_render(self._ctx, xyz.ctypes.data_as(L._fp),
                                        ex.ctypes.data_as(L._u8p)))
        return xyz, ex

    def diagnostics(self):
        d = L.NbDiag()
        self._check(lib.nb_diagnostics(self._ctx, C.byref(d)))
        return {"kinetic": d.kinetic, "potential": d.potential,
                "momentum": np.array(list(d.momentum)), "mass": d.mass,
                "n_active": d.n_active}

    def set_timing(self, on=True):
        self._check(lib.nb_set_timing(self._ctx, 1 if on else 0))

    def config(self):
        cfg = (C.c_int32 * 5)()
        self._check(lib.nb_get_config(self._ctx, cfg))
        return dict(zip(("ipl", "tile_j", "jsplit", "block", "variant"), list(cfg)))

    def autotune(self, reps=3):
        cfg = (C.c_int32 * 5)()
        self._check(lib.nb_autotune(self._ctx, int(reps), cfg))
        return dict(zip(("ipl", "tile_j", "jsplit", "block", "variant"), list(cfg)))

    # ---- learned far-field surrogate ----------------------------------------------------------
    def sur_enable(self, grid=32, shell=2, r0=0.0, use_mlp=True, check_every=1,
                   check_samples=256, tol=1e-3, seed=12345, tol_body=0.0, body_quantile=0.99):
        """Replace the O(N^2) force phase by the cell-grid multipole + MLP surrogate (FP32 path,
        eps > 0, no overlap cutoff; with collisions every radius must be <= r0 / grid).  Steps
        whose sampled error exceeds tol, or (tol_body > 0) whose per-body errors exceed tol_body
        on more than a fraction 1 - body_quantile of the samples, are redone by direct
        summation."""
        p = L.NbSurParams()
        lib.nb_sur_params_init(C.byref(p))
        p.grid, p.shell, p.r0 = int(grid), int(shell), float(r0)
        p.use_mlp, p.check_every = int(bool(use_mlp)), int(check_every)
        p.check_samples, p.tol, p.seed = int(check_samples), float(tol), int(seed)
        p.tol_body, p.body_quantile = float(tol_body), float(body_quantile)
        self._check(lib.nb_sur_enable(self._ctx, C.byref(p)))
        return self

    def sur_disable(self):
        self._check(lib.nb_sur_disable(self._ctx))

    def sur_collect(self, samples, seed=0):
        """Append training rows from the loaded state, labelled by the FP32 direct sum."""
        self._check(lib.nb_sur_collect(self._ctx, int(samples), int(seed)))

    def sur_fit(self, steps=2000, batch=256, lr=1e-3):
        loss = C.c_float()
        self._check(lib.nb_sur_fit(self._ctx, int(steps), int(batch), float(lr), C.byref(loss)))
        return loss.value

    def sur_error(self, samples=1024, use_mlp=True):
        err = C.c_double()
        self._check(lib.nb_sur_error(self._ctx, int(samples), int(bool(use_mlp)), C.byref(err)))
        return err.value

    def sur_stats(self):
        st = L.NbSurStats()
        self._check(lib.nb_sur_get_stats(self._ctx, C.byref(st)))
        return st.as_dict()

    def sur_save(self, path):
        self._check(lib.nb_sur_save(self._ctx, str(path).encode()))

    def sur_load(self, path):
        self._check(lib.nb_sur_load(self._ctx, str(path).encode()))


class NBodyBatch:
    """Many independent simulations in one persistent, graph-captured launch (nb_batch_*)."""

    def __init__(self, G=1.0, eps=0.01, device=0):
        p = L.NbParams()
        check(lib.nb_params_init(C.byref(p)))
        p.G, p.eps, p.device = float(G), float(eps), int(device)
        p.overlap_cutoff = 0
        p.collisions = 0
        self._b = C.c_void_p()
        rc = lib.nb_batch_create(C.byref(p), C.byref(self._b))
        if rc != L.NB_OK:
            raise L.NBodyError(rc, "nb_batch_create failed")
        self.offsets = None

    def _check(self, rc):
        if rc != L.NB_OK:
            raise L.NBodyError(rc, lib.nb_batch_last_error(self._b).decode())

    def close(self):
        if self._b:
            lib.nb_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, sims):
        """sims: list of Bodies."""
        from .gen import concat
        b = concat(*sims)
        off = np.zeros(len(sims) + 1, np.int64)
        off[1:] = np.cumsum([s.n for s in sims])
        self.offsets = off
        self.n = b.n
        self._check(lib.nb_batch_load(self._b, len(sims), off.ctypes.data_as(L._i64p), _dp(b.x),
                                      _dp(b.y), _dp(b.z), _dp(b.vx), _dp(b.vy), _dp(b.vz),
                                      _dp(b.m), b.flags.ctypes.data_as(L._u8p)))
        self._all = b
        return self

    def step(self, dt, R=1.0, nsteps=1):
        st = L.NbStats()
        self._check(lib.nb_batch_step(self._b, float(dt), float(R), int(nsteps), C.byref(st)))
        return st.as_dict()

    def set_graph(self, on):
        self._check(lib.nb_batch_set_graph(self._b, 1 if on else 0))

    def forces(self):
        f = np.empty((3, self.n), np.float64)
        self._check(lib.nb_batch_forces(self._b, _dp(f[0]), _dp(f[1]), _dp(f[2])))
        return f.T.copy()

    def store(self):
        b = self._all.copy()
        self._check(lib.nb_batch_store(self._b, _dp(b.x), _dp(b.y), _dp(b.z), _dp(b.vx),
                                       _dp(b.vy), _dp(b.vz), b.flags.ctypes.data_as(L._u8p)))
        return b

    def shapes(self):
        s = (C.c_int32 * 3)()
        self._check(lib.nb_batch_get_shapes(self._b, s))
        return list(s)

    def set_shapes(self, ipl3):
        s = (C.c_int32 * 3)(*ipl3)
        self._check(lib.nb_batch_set_shapes(self._b, s))

    def autotune(self, reps=2):
        """Tune waves per item, and IPL and Kahan tile per size bucket; returns the tuning."""
        s = (C.c_int32 * 3)()
        self._check(lib.nb_batch_autotune(self._b, int(reps), s))
        return self.tuning()

    def tuning(self):
        """{"waves": W, "buckets": [(ipl, tile) for N < 1024, < 4096, >= 4096]}"""
        c = (C.c_int32 * 7)()
        self._check(lib.nb_batch_get_tuning(self._b, c))
        return {"waves": c[0], "buckets": [(c[1 + 2 * k], c[2 + 2 * k]) for k in range(3)]}

    def set_tuning(self, waves, buckets):
        c = (C.c_int32 * 7)(int(waves), *[int(v) for ipl_tile in buckets for v in ipl_tile])
        self._check(lib.nb_batch_set_tuning(self._b, c))
