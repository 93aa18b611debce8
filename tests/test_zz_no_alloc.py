"""nb_forces_noalloc: the force phase with its intermediates in a fixed device array, called
through the C-ABI on device buffers, against nb_compute_forces of a context with the same G, eps
and kernel shape (bit for bit), on two calls in a row, the second smaller."""
import ctypes as C
import os

import numpy as np
import pytest

from nbodygo_amd import NBodyGPU, gen
from nbodygo_amd import _lib as L

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


class _Hip:
    """The few HIP runtime calls the test needs to hold inputs and outputs on the device."""

    def __init__(self):
        self.lib = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib",
                                       "libamdhip64.so"))
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipDeviceSynchronize.argtypes = []

    def ok(self, rc, what):
        assert rc == 0, "%s returned %d" % (what, rc)

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self.ok(self.lib.hipMalloc(C.byref(p), max(a.nbytes, 1)), "hipMalloc")
        self.ok(self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, _H2D), "hipMemcpy")
        return p

    def empty(self, nbytes):
        p = C.c_void_p()
        self.ok(self.lib.hipMalloc(C.byref(p), max(nbytes, 1)), "hipMalloc")
        return p

    def get(self, p, n, dtype):
        out = np.empty(n, dtype)
        self.ok(self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, _D2H), "hipMemcpy")
        return out

    def free(self, *ps):
        for p in ps:
            self.ok(self.lib.hipFree(p), "hipFree")


def _bodies(n, seed):
    b = gen.plummer(n, seed=seed)
    rng = np.random.default_rng(seed)
    off = rng.choice(n, n // 40, replace=False)
    b.flags[off[: len(off) // 2]] = 0                       # removed
    b.flags[off[len(off) // 2:]] = L.NB_F_EXISTS | L.NB_F_FRAGMENTING
    return b


def _context_forces(b, G, eps):
    npad = -(-b.n // 1024) * 1024
    with NBodyGPU(G=G, eps=eps, overlap_cutoff=False, collisions=False, ipl=1, tile_j=512,
                  jsplit=min(4, npad // 512), block=256, variant=2) as g:
        g.load(b)
        return g.forces()


def _noalloc_forces(hip, b, G, eps, with_flags=True):
    ins = [hip.put(getattr(b, k)) for k in ("x", "y", "z", "m")]
    fl = hip.put(b.flags) if with_flags else None
    outs = [hip.empty(8 * b.n) for _ in range(3)]
    rc = L.lib.nb_forces_noalloc(0, b.n, *ins, fl, G, eps, *outs, None)
    msg = L.lib.nb_last_error(None)
    assert rc == L.NB_OK, (rc, msg)
    hip.ok(hip.lib.hipDeviceSynchronize(), "hipDeviceSynchronize")
    f = np.stack([hip.get(p, b.n, np.float64) for p in outs], axis=1)
    hip.free(*ins, *outs)
    if fl is not None:
        hip.free(fl)
    return f


def test_noalloc_matches_compute_forces_on_two_calls():
    hip = _Hip()
    G, eps = 1.0, 0.01
    for n, seed in ((20000, 11), (3000, 12)):  # the second call is smaller
        print("nb_forces_noalloc: n=%d seed=%d" % (n, seed), flush=True)
        b = _bodies(n, seed)
        got = _noalloc_forces(hip, b, G, eps)
        exp = _context_forces(b, G, eps)
        np.testing.assert_array_equal(got, exp)
        act = b.active()
        assert np.all(np.isfinite(got[act])) and np.abs(got[act]).max() > 0


def test_noalloc_without_flags_and_at_the_limit_of_one_tile_pair():
    hip = _Hip()
    b = gen.plummer(1000, seed=5)          # one 1024 pad, two 512-source tiles, two slices
    got = _noalloc_forces(hip, b, 2.5, 0.05, with_flags=False)
    exp = _context_forces(b, 2.5, 0.05)
    np.testing.assert_array_equal(got, exp)


def test_noalloc_rejects_sizes_beyond_its_scratch():
    n = L.NB_NOALLOC_MAX_BODIES + 1
    rc = L.lib.nb_forces_noalloc(0, n, *([C.c_void_p(16)] * 4), None, 1.0, 0.01,
                                 *([C.c_void_p(16)] * 3), None)
    assert rc == L.NB_E_INVALID
    assert b"NB_NOALLOC_MAX_BODIES" in L.lib.nb_last_error(None)
    rc = L.lib.nb_forces_noalloc(0, 10, *([C.c_void_p(16)] * 4), None, 1.0, 0.0,
                                 *([C.c_void_p(16)] * 3), None)
    assert rc == L.NB_E_INVALID
