"""The collision pass with pairs that share no body resolved together (nb_params.resolve = 0)
against the one-lane pass in canonical order (resolve = 1), bit for bit, on dense scenes whose
pair lists need the multi-workgroup sort; and the pair lists against the oracle's detection.
The reference resolves its events one at a time (ProcessMods, body_collection.go:212-233)."""
import numpy as np
import pytest

from nbodygo_amd import NBodyGPU, gen

pytestmark = pytest.mark.gpu

G_REF = 6.673e-11


def _run(b, resolve, steps, **kw):
    """`steps` cycles one call each: (final state, per-cycle stats, pairs of the last cycle)."""
    with NBodyGPU(resolve=resolve, pair_capacity=1 << 20, **kw) as g:
        g.load(b)
        stats = [g.step(1e-3, 1.0) for _ in range(steps)]
        return g.store(), stats, g.pairs()


def _same_state(s1, s2):
    for k in ("x", "y", "z", "vx", "vy", "vz", "m", "rest"):
        np.testing.assert_array_equal(getattr(s1, k), getattr(s2, k), err_msg=k)
    np.testing.assert_array_equal(s1.flags, s2.flags)


@pytest.mark.parametrize("n,radius_k,steps", [(3000, 0.25, 3), (1000, 0.05, 30), (20000, 0.12, 2)])
def test_parallel_resolution_is_bit_identical_to_one_lane(oracle, n, radius_k, steps):
    b = gen.uniform_sphere(n, seed=1, radius_k=radius_k)
    print("resolve parity: n=%d radius_k=%g steps=%d" % (n, radius_k, steps), flush=True)
    par, sp, pp = _run(b, 0, steps)
    ser, ss, ps = _run(b, 1, steps)
    _same_state(par, ser)
    np.testing.assert_array_equal(pp, ps)
    for a, c in zip(sp, ss):
        for k in ("pairs_total", "resolved_total", "fragment_requests", "removed_total"):
            assert a[k] == c[k], k
    assert sum(s["resolved_total"] for s in sp) > 0


def test_dense_pairs_sorted_like_the_oracle_and_resolution_faster(oracle):
    """About 350,000 pairs in one cycle: the multi-workgroup sort gives the oracle's canonical
    list, the states agree bit for bit, and the parallel pass (HIP-event time of the resolution
    kernels alone) takes under a tenth of the one-lane pass's time."""
    b = gen.uniform_sphere(20000, seed=2, radius_k=0.12)
    st = oracle.State.from_bodies(b)
    ref_pairs = oracle.detect(st)
    assert len(ref_pairs) > 30000
    times = {}
    states = {}
    for resolve in (0, 1):
        with NBodyGPU(resolve=resolve, pair_capacity=1 << 20) as g:
            g.load(b)
            g.forces()
            np.testing.assert_array_equal(g.pairs(), ref_pairs)
            g.load(b)
            g.set_timing(True)
            s = g.step(1e-3, 1.0)
            times[resolve] = s["resolve_ms"]
            states[resolve] = g.store()
    _same_state(states[0], states[1])
    print("one cycle with %d pairs: resolution parallel %.3f ms, one lane %.3f ms"
          % (len(ref_pairs), times[0], times[1]), flush=True)
    assert 0 < times[0] * 10 < times[1], times
