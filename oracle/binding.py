"""ctypes binding of the C oracle (oracle/nb_oracle.h) -- TEST INFRASTRUCTURE ONLY.

Works on plain numpy arrays (no dependency on the product package), so the oracle can check the
GPU path without sharing code with it.
"""
import ctypes as C
import os
import subprocess
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_nb.so")

F_EXISTS, F_FRAGMENTING, F_SUN, F_PINNED = 1, 2, 4, 8
CB_NONE, CB_SUBSUME, CB_ELASTIC, CB_FRAGMENT = 0, 1, 2, 3


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


class OrState(C.Structure):
    _fields_ = [("n", C.c_int64)] + [(k, C.POINTER(C.c_double)) for k in
                                     ("x", "y", "z", "vx", "vy", "vz", "m", "r", "rest",
                                      "frag")] + [
        ("beh", C.POINTER(C.c_int32)), ("flags", C.POINTER(C.c_uint8)),
        ("collided", C.POINTER(C.c_uint8))]


class OrParams(C.Structure):
    _fields_ = [("G", C.c_double), ("eps", C.c_double), ("cutoff", C.c_int32),
                ("collisions", C.c_int32), ("nthreads", C.c_int32)]


class OrStats(C.Structure):
    _fields_ = [("pairs_total", C.c_int64), ("pairs_last", C.c_int64),
                ("resolved_total", C.c_int64), ("fragment_requests", C.c_int64),
                ("removed_total", C.c_int64)]


class OrBhParams(C.Structure):
    _fields_ = [("G", C.c_double), ("theta", C.c_double), ("soft", C.c_double),
                ("max_depth", C.c_int32), ("pad", C.c_int32)]


class OrBhInfo(C.Structure):
    _fields_ = [("nodes", C.c_int64), ("internal", C.c_int64), ("buckets", C.c_int64),
                ("interactions", C.c_int64), ("max_depth_seen", C.c_int32), ("pad", C.c_int32),
                ("root_mass", C.c_double), ("root_half", C.c_double),
                ("root_cm", C.c_double * 3)]


_dp = C.POINTER(C.c_double)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        sp = C.POINTER(OrState)
        L.or_force_ref.argtypes = [sp, C.c_double, C.POINTER(C.c_int64), C.c_int64, _dp, _dp, _dp,
                                   C.c_int]
        L.or_force_soft.argtypes = [sp, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_int64),
                                    C.c_int64, _dp, _dp, _dp, C.c_int]
        L.or_force_soft_ld.argtypes = [sp, C.c_double, C.c_double, C.c_int,
                                       C.POINTER(C.c_int64), C.c_int64, _dp, _dp, _dp]
        L.or_detect.argtypes = [sp, C.POINTER(C.c_int32), C.c_int64]
        L.or_detect.restype = C.c_int64
        L.or_detect_grid.argtypes = [sp, C.POINTER(C.c_int32), C.c_int64, C.c_int]
        L.or_detect_grid.restype = C.c_int64
        L.or_elastic.argtypes = [sp, C.c_int64, C.c_int64, _dp]
        L.or_resolve.argtypes = [sp, C.POINTER(C.c_int32), C.c_int64, C.POINTER(OrStats)]
        L.or_update.argtypes = [sp, _dp, _dp, _dp, C.c_double, C.c_double, C.POINTER(OrStats)]
        L.or_step.argtypes = [sp, C.POINTER(OrParams), C.c_double, C.c_double, C.c_int,
                              C.POINTER(C.c_int32), C.c_int64, C.POINTER(OrStats)]
        L.or_energy.argtypes = [sp, C.c_double, C.c_double, _dp, C.c_int]
        L.or_bench_workpool.argtypes = [sp, C.c_double, C.c_int64, C.c_int, C.c_int, _dp,
                                        C.POINTER(C.c_int64)]
        L.or_bench_workpool.restype = C.c_double
        L.or_bench_cycles.argtypes = [sp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                      C.POINTER(OrStats)]
        L.or_bench_cycles.restype = C.c_double
        L.or_bh_force.argtypes = [sp, C.POINTER(OrBhParams), _dp, _dp, _dp,
                                  C.POINTER(C.c_int32), C.POINTER(OrBhInfo)]
        L.or_bh_force.restype = C.c_int
        L.or_soa_soft.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                  C.c_int64, C.c_int64, _dp, _dp, _dp, C.c_int]
        L.or_soa_soft.restype = None
        _lib = L
    return _lib


FIELDS = ("x", "y", "z", "vx", "vy", "vz", "m", "r")


class State:
    """Owns contiguous copies of a body set; .s is the or_state view."""

    def __init__(self, x, y, z, vx, vy, vz, m, r, beh=None, flags=None, rest=None, frag=None):
        n = len(x)
        self.a = {}
        for k, v in zip(FIELDS, (x, y, z, vx, vy, vz, m, r)):
            self.a[k] = np.array(v, dtype=np.float64, copy=True)
        self.a["rest"] = np.ones(n) if rest is None else np.array(rest, np.float64, copy=True)
        self.a["frag"] = np.zeros(n) if frag is None else np.array(frag, np.float64, copy=True)
        self.beh = (np.full(n, CB_ELASTIC, np.int32) if beh is None
                    else np.array(beh, np.int32, copy=True))
        self.flags = (np.full(n, F_EXISTS, np.uint8) if flags is None
                      else np.array(flags, np.uint8, copy=True))
        self.collided = np.zeros(n, np.uint8)
        self.n = n
        s = OrState()
        s.n = n
        for k in FIELDS + ("rest", "frag"):
            setattr(s, k, self.a[k].ctypes.data_as(_dp))
        s.beh = self.beh.ctypes.data_as(C.POINTER(C.c_int32))
        s.flags = self.flags.ctypes.data_as(C.POINTER(C.c_uint8))
        s.collided = self.collided.ctypes.data_as(C.POINTER(C.c_uint8))
        self.s = s

    @classmethod
    def from_bodies(cls, b):
        return cls(b.x, b.y, b.z, b.vx, b.vy, b.vz, b.m, b.r, b.beh, b.flags, b.rest, b.frag)

    def __getattr__(self, k):
        a = self.__dict__.get("a")
        if a is not None and k in a:
            return a[k]
        raise AttributeError(k)

    def pos(self):
        return np.stack([self.a["x"], self.a["y"], self.a["z"]], axis=1)

    def vel(self):
        return np.stack([self.a["vx"], self.a["vy"], self.a["vz"]], axis=1)


def _idx(idx):
    if idx is None:
        return None, 0, None
    idx = np.ascontiguousarray(idx, np.int64)
    return idx, idx.shape[0], idx.ctypes.data_as(C.POINTER(C.c_int64))


def force_ref(st, G=6.673e-11, idx=None, nthreads=8):
    idx, k, ip = _idx(idx)
    k = st.n if idx is None else k
    f = np.zeros((3, k))
    lib().or_force_ref(C.byref(st.s), G, ip, k, *(f[c].ctypes.data_as(_dp) for c in range(3)),
                       nthreads)
    return f.T.copy()


def force_soft(st, G, eps, cutoff=False, idx=None, nthreads=8, long_double=False):
    idx, k, ip = _idx(idx)
    k = st.n if idx is None else k
    f = np.zeros((3, k))
    if long_double:
        lib().or_force_soft_ld(C.byref(st.s), G, eps, int(cutoff), ip, k,
                               *(f[c].ctypes.data_as(_dp) for c in range(3)))
    else:
        lib().or_force_soft(C.byref(st.s), G, eps, int(cutoff), ip, k,
                            *(f[c].ctypes.data_as(_dp) for c in range(3)), nthreads)
    return f.T.copy()


def soa_soft(x, y, z, m, G, eps, t0=0, t1=None, nthreads=8):
    """or_soa_soft: the benchmark's optimised CPU line (vectorised, threaded FP64 softened sum
    over SoA arrays) for targets [t0, t1); returns (accelerations [k x 3], seconds)."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, m)]
    n = len(arrs[0])
    t1 = n if t1 is None else t1
    out = np.zeros((3, t1 - t0))
    t = time.perf_counter()
    lib().or_soa_soft(n, *(a.ctypes.data_as(_dp) for a in arrs), G, eps, t0, t1,
                      *(out[c].ctypes.data_as(_dp) for c in range(3)), nthreads)
    return out.T.copy(), time.perf_counter() - t


def detect(st):
    n = lib().or_detect(C.byref(st.s), None, 0)
    out = np.zeros((max(n, 1), 2), np.int32)
    lib().or_detect(C.byref(st.s), out.ctypes.data_as(C.POINTER(C.c_int32)), n)
    return out[:n]


def detect_grid(st, nthreads=8):
    """or_detect_grid: the canonical pair list by cell lists (large N)."""
    n = lib().or_detect_grid(C.byref(st.s), None, 0, nthreads)
    if n < 0:
        raise RuntimeError("or_detect_grid failed %d" % n)
    out = np.zeros((max(n, 1), 2), np.int32)
    lib().or_detect_grid(C.byref(st.s), out.ctypes.data_as(C.POINTER(C.c_int32)), n, nthreads)
    return out[:n]


def elastic(st, i, j):
    out = np.zeros(9)
    ok = lib().or_elastic(C.byref(st.s), i, j, out.ctypes.data_as(_dp))
    return bool(ok), out


def resolve(st, pairs):
    pairs = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
    stats = OrStats()
    lib().or_resolve(C.byref(st.s), pairs.ctypes.data_as(C.POINTER(C.c_int32)), pairs.shape[0],
                     C.byref(stats))
    return {f[0]: getattr(stats, f[0]) for f in OrStats._fields_}


def update(st, f, dt, R):
    f = np.ascontiguousarray(np.asarray(f, np.float64).T)
    stats = OrStats()
    lib().or_update(C.byref(st.s), *(f[c].ctypes.data_as(_dp) for c in range(3)), dt, R,
                    C.byref(stats))
    return {k[0]: getattr(stats, k[0]) for k in OrStats._fields_}


def step(st, G, eps, cutoff, collisions, dt, R, nsteps=1, nthreads=8):
    p = OrParams(G, eps, int(cutoff), int(collisions), nthreads)
    stats = OrStats()
    cap = 1 << 20
    pairs = np.zeros((cap, 2), np.int32)
    rc = lib().or_step(C.byref(st.s), C.byref(p), dt, R, nsteps,
                       pairs.ctypes.data_as(C.POINTER(C.c_int32)), cap, C.byref(stats))
    if rc:
        raise RuntimeError("or_step failed %d" % rc)
    d = {k[0]: getattr(stats, k[0]) for k in OrStats._fields_}
    d["pairs"] = pairs[:min(d["pairs_last"], cap)].copy()
    return d


def energy(st, G, eps, nthreads=8):
    out = np.zeros(5)
    lib().or_energy(C.byref(st.s), G, eps, out.ctypes.data_as(_dp), nthreads)
    return {"kinetic": out[0], "potential": out[1], "momentum": out[2:5].copy()}


def bench_workpool(st, G, ntargets, nsteps=1, nthreads=8, want_force=False):
    f = np.zeros((ntargets, 3)) if want_force else None
    hits = C.c_int64()
    el = lib().or_bench_workpool(C.byref(st.s), G, ntargets, nsteps, nthreads,
                                 f.ctypes.data_as(_dp) if want_force else None, C.byref(hits))
    return el, hits.value, f


def bench_cycles(st, G, dt, R, nsteps, nthreads=8):
    """nsteps whole cycles of the reference's CPU algorithm (or_bench_cycles) on st, in place:
    (elapsed seconds, stats dict)."""
    stats = OrStats()
    el = lib().or_bench_cycles(C.byref(st.s), G, dt, R, nsteps, nthreads, C.byref(stats))
    if el < 0:
        raise RuntimeError("or_bench_cycles failed")
    return el, {k[0]: getattr(stats, k[0]) for k in OrStats._fields_}


def bh_force(st, G=6.673e-11, theta=0.3, soft=1e-4, max_depth=21):
    """Barnes-Hut force (or_bh_force): (n x 3 forces, per-body leaf depth, tree info)."""
    p = OrBhParams(G, theta, soft, max_depth, 0)
    f = np.zeros((3, st.n))
    depth = np.zeros(max(st.n, 1), np.int32)
    info = OrBhInfo()
    rc = lib().or_bh_force(C.byref(st.s), C.byref(p), *(f[c].ctypes.data_as(_dp)
                                                        for c in range(3)),
                           depth.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(info))
    if rc:
        raise RuntimeError("or_bh_force failed %d" % rc)
    d = {k[0]: getattr(info, k[0]) for k in OrBhInfo._fields_ if k[0] != "pad"}
    d["root_cm"] = np.array(list(info.root_cm))
    return f.T.copy(), depth[:st.n].copy(), d
