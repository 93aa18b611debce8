"""GPU tests of the device collection lifecycle: nb_remove_bodies (processDeletes,
computation-runner.go:176-216), nb_add_bodies + nb_cycle (BodyCollection.Cycle,
body_collection.go:253-296), nb_get_body / nb_set_body (GetBody / ModBody,
body_collection.go:111-191) and nb_store_bodies.

Bars, all bit-exact: removal sets, flags and masses equal the lifecycle oracle's walk; the state
after nb_cycle equals the oracle's compaction + adds; forces and later steps after a cycle or a
patch equal those of a fresh context loaded with the same arrays (so every derived buffer -- the
FP32 source stream, per-tile radii, pair buffers, the active count, the largest radius, the
surrogate's per-body buffers -- was rebuilt).
"""
import numpy as np
import pytest

from nbodygo_amd import NBodyGPU, NBodyError, gen
from nbodygo_amd import _lib as L
from nbodygo_amd.sim import Bodies
from oracle import lifecycle as LC

pytestmark = pytest.mark.gpu

EPS = 0.01
KEYS = ("x", "y", "z", "vx", "vy", "vz", "m", "r", "beh", "flags", "rest", "frag")


def sim(**kw):
    kw.setdefault("collisions", False)
    return NBodyGPU(G=1.0, eps=EPS, overlap_cutoff=False, **kw)


def mixed_flags(b, rng, p_gone=0.1, p_pin=0.1, p_frag=0.03):
    f = np.full(b.n, L.NB_F_EXISTS, np.uint8)
    f[rng.random(b.n) < p_gone] = 0
    f[rng.random(b.n) < p_pin] |= L.NB_F_PINNED
    f[rng.random(b.n) < p_frag] |= L.NB_F_FRAGMENTING
    b.flags[:] = f
    return b


def assert_same(a, b, keys=KEYS):
    for k in keys:
        np.testing.assert_array_equal(getattr(a, k), b[k] if isinstance(b, dict) else
                                      getattr(b, k), err_msg=k)


def n_active(g):
    return g.step(0.0, nsteps=0)["n_active"]


def test_remove_walk_matches_oracle():
    rng = np.random.default_rng(1)
    cases = [(10, 3), (7, 2), (9, 3), (1, 1), (5, 50), (1000, 1000), (1000, 999), (4099, 7),
             (4099, -1), (65536, 1), (100000, 1000), (100000, 33333), (250001, 4096)]
    cases += [(int(rng.integers(1, 5000)), int(rng.integers(1, 6000))) for _ in range(20)]
    with sim() as g:
        for c, (n, d) in enumerate(cases):
            b = mixed_flags(gen.uniform_sphere(n, seed=c), rng,
                            p_gone=rng.uniform(0, 0.6), p_pin=rng.uniform(0, 0.5))
            g.load(b)
            k = g.remove_bodies(d)
            ex = (b.flags & L.NB_F_EXISTS).astype(bool)
            idx = LC.remove_indices(ex, (b.flags & L.NB_F_PINNED).astype(bool), d)
            exp = b.copy()
            LC.apply_removals(exp, idx)
            s = g.store()
            assert k == len(idx), (n, d)
            np.testing.assert_array_equal(s.flags, exp.flags, err_msg=str((n, d)))
            np.testing.assert_array_equal(s.m, exp.m, err_msg=str((n, d)))
            assert n_active(g) == int(exp.active().sum())
    with sim() as g:
        g.load(gen.uniform_sphere(100, seed=9))
        assert g.remove_bodies(0) == 0
        with pytest.raises(NBodyError) as e:
            g.remove_bodies(-2)
        assert e.value.code == L.NB_E_INVALID


@pytest.mark.parametrize("collisions", [False, True])
def test_cycle_matches_oracle_and_a_fresh_load(collisions):
    rng = np.random.default_rng(2)
    b = mixed_flags(gen.plummer(20000, seed=2), rng, p_gone=0.05, p_pin=0.2, p_frag=0.0)
    b.r[:] = 2e-3
    b.beh[:] = rng.integers(0, 4, b.n)
    b.frag[:] = rng.random(b.n)
    adds = [gen.plummer(3000, seed=3), gen.plummer(2500, seed=4)]  # grows past the capacity
    for a in adds:
        a.r[:] = 2e-3
        a.flags[::17] |= L.NB_F_SUN
    with sim(collisions=collisions) as g:
        g.load(b)
        g.step(1e-4, 1.0, 2)
        assert g.remove_bodies(777) >= 777
        before = g.store()
        for a in adds:
            g.add_bodies(a)
        assert g.cycle(0.7) == int((before.flags & L.NB_F_EXISTS).astype(bool).sum()) + 5500
        exp = LC.cycle(before, adds, 0.7)
        after = g.store()
        assert_same(after, exp)
        f_cycled = g.forces()
        g.step(1e-4, 0.7, 3)
        stepped = g.store()
        pairs = g.pairs()
        na = n_active(g)
        # a second cycle that shrinks the collection
        g.remove_bodies(5000)
        mid = g.store()
        g.cycle(0.7)
        shrunk = g.store()
        f_shrunk = g.forces()
    assert_same(shrunk, LC.cycle(mid, [], 0.7))
    with sim(collisions=collisions) as h:
        h.load(Bodies(*(exp[k] for k in KEYS[:8]), beh=exp["beh"], flags=exp["flags"],
                      rest=exp["rest"], frag=exp["frag"]))
        np.testing.assert_array_equal(h.forces(), f_cycled)
        h.step(1e-4, 0.7, 3)
        assert_same(h.store(), stepped)
        np.testing.assert_array_equal(h.pairs(), pairs)
        assert n_active(h) == na
        h.load(shrunk)
        np.testing.assert_array_equal(h.forces(), f_shrunk)


def test_empty_collection_adds_and_remove_all():
    with sim() as g:
        g.load(gen.uniform_sphere(0, seed=1))
        assert g.cycle(1.0) == 0
        a = gen.uniform_sphere(1500, seed=5)
        g.add_bodies(a)
        assert g.cycle(0.5) == 1500
        exp = LC.cycle(gen.uniform_sphere(0, seed=1), [a], 0.5)
        assert_same(g.store(), exp)
        f = g.forces()
        assert g.remove_bodies(-1) == 1500
        assert n_active(g) == 0
        assert g.cycle(1.0) == 0
        g.step(1e-3, 1.0, 1)
    with sim() as h:
        h.load(Bodies(*(exp[k] for k in KEYS[:8]), beh=exp["beh"], flags=exp["flags"],
                      rest=exp["rest"], frag=exp["frag"]))
        np.testing.assert_array_equal(h.forces(), f)


def test_get_and_set_body():
    b = gen.plummer(8192, seed=6)
    with sim() as g:
        g.load(b)
        d = g.get_body(17)
        for k in KEYS:
            assert d[k] == getattr(b, k)[17], k
        with pytest.raises(NBodyError):
            g.get_body(8192)
        with pytest.raises(NBodyError):
            g.get_body(-1)
        g.set_body(17, x=0.25, vx=-1.5, m=3.0 * b.m[17], r=0.05)
        g.set_body(100, flags=0)                           # stops existing
        g.set_body(200, flags=L.NB_F_EXISTS | L.NB_F_FRAGMENTING)
        mod = b.copy()
        mod.x[17], mod.vx[17], mod.m[17], mod.r[17] = 0.25, -1.5, 3.0 * b.m[17], 0.05
        mod.flags[100] = 0
        mod.flags[200] = L.NB_F_EXISTS | L.NB_F_FRAGMENTING
        assert_same(g.store(), mod)
        assert n_active(g) == int(mod.active().sum())
        f = g.forces()
        g.step(1e-3, 1.0, 2)
        s = g.store()
    with sim() as h:
        h.load(mod)
        np.testing.assert_array_equal(h.forces(), f)
        h.step(1e-3, 1.0, 2)
        assert_same(h.store(), s)


def test_surrogate_after_cycle_matches_a_fresh_context():
    b = gen.plummer(32768, seed=7)
    kw = dict(grid=32, shell=2, r0=0.4, use_mlp=False, check_every=0)
    with sim() as g:
        g.load(b)
        g.sur_enable(**kw)
        g.forces()
        g.remove_bodies(3000)
        g.add_bodies(gen.plummer(9000, seed=8))
        g.cycle(1.0)
        state = g.store()
        f = g.forces()
    with sim() as h:
        h.load(state)
        h.sur_enable(**kw)
        np.testing.assert_array_equal(h.forces(), f)
