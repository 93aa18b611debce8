"""GPU tests of the Barnes-Hut octree (nb_bh_*) against the octree oracle
(oracle/nb_oracle_bh.c: cmd/body/octree.go with Build wired into BuildTree).

Bars:
  topology     each body's leaf depth, the number of split cells and of buckets equal the
               oracle's (integers, exact)
  theta = 0    every leaf visited in the reference's order with the reference's arithmetic:
               forces bit-identical to the oracle
  theta = 0.3  node centres of mass are summed bottom-up (the reference folds bodies in array
               order): per-body relative force difference <= 1e-11
  suns         direct sums in array order: bit-identical
  collisions   pair sets identical to the oracle's; a step matches the oracle's force + detect +
               resolve + update (1e-12: coll3d's transcendentals)
"""
import numpy as np
import pytest

from nbodygo_amd import NBodyGPU, NBodyError, gen
from nbodygo_amd import _lib as L

pytestmark = pytest.mark.gpu


def run(b, theta=0.3, max_depth=21, collisions=False, term=0):
    with NBodyGPU(G=1.0, precision="fp64", collisions=collisions) as g:
        g.load(b)
        g.bh_enable(theta=theta, max_depth=max_depth, term=term)
        f = g.forces()
        return f, g.bh_depths(), g.bh_stats(), (g.pairs() if collisions else None)


def rel_diff(a, b):
    na = np.linalg.norm(b, axis=1)
    return np.linalg.norm(a - b, axis=1) / np.where(na > 0, na, 1.0)


def mixed(b, seed):
    rng = np.random.default_rng(seed)
    b.flags[rng.random(b.n) < 0.05] = 0
    b.flags[rng.random(b.n) < 0.02] |= L.NB_F_FRAGMENTING
    return b


@pytest.mark.parametrize("term", [1, 2])  # three divisions; one shared reciprocal
def test_topology_and_theta_zero_are_exact(oracle, term):
    b = mixed(gen.plummer(8000, seed=1), 1)
    st = oracle.State.from_bodies(b)
    f, depth, stats, _ = run(b, theta=0.0, term=term)
    fo, do, info = oracle.bh_force(st, G=1.0, theta=0.0)
    np.testing.assert_array_equal(depth, do)
    assert stats["nodes"] - stats["buckets"] == info["internal"]
    assert stats["buckets"] == info["buckets"] == 0
    assert stats["depth_max"] == info["max_depth_seen"]
    assert stats["interactions"] == info["interactions"]
    assert stats["overflow"] == 0
    np.testing.assert_array_equal(f, fo)


@pytest.mark.parametrize("kind,n", [("plummer", 65536), ("two_plummer", 30000),
                                    ("uniform_sphere", 30000)])
def test_default_theta_matches_oracle(oracle, kind, n):
    b = getattr(gen, kind)(n, seed=2)
    st = oracle.State.from_bodies(b)
    f, depth, stats, _ = run(b)
    fo, do, info = oracle.bh_force(st, G=1.0)
    np.testing.assert_array_equal(depth, do)
    assert stats["nodes"] - stats["buckets"] == info["internal"]
    d = rel_diff(f, fo)
    assert d.max() <= 1e-11, d.max()
    assert abs(stats["interactions"] - info["interactions"]) <= 1e-6 * info["interactions"]
    # and it is a Barnes-Hut force: within the opening-angle error of the direct sum
    ref = oracle.force_soft(st, 1.0, 1e-4, idx=np.arange(0, b.n, 97), nthreads=16)
    assert np.linalg.norm(f[::97] - ref) / np.linalg.norm(ref) < 3e-3


@pytest.mark.parametrize("n", [65536, 262144])
def test_the_two_term_forms_are_bit_identical(n):
    """The exact walk's term with three divisions and with one shared reciprocal give the same
    bits at theta 0.3 (at 262,144 bodies the default picks the shared form, at 65,536 the
    divisions)."""
    b = gen.plummer(n, seed=21)
    f1 = run(b, term=1)[0]
    f2 = run(b, term=2)[0]
    f0 = run(b)[0]
    np.testing.assert_array_equal(f1, f2)
    np.testing.assert_array_equal(f0, f1)


def test_suns_are_direct_sums_bit_exact(oracle):
    b = gen.plummer(5000, seed=3)
    b.flags[[0, 77, 4999]] |= L.NB_F_SUN
    b.m[[0, 77, 4999]] *= 1e3
    st = oracle.State.from_bodies(b)
    f, _, _, _ = run(b)
    fo, _, _ = oracle.bh_force(st, G=1.0)
    np.testing.assert_array_equal(f[[0, 77, 4999]], fo[[0, 77, 4999]])
    assert rel_diff(f, fo).max() <= 1e-11


def test_coincident_bodies_and_the_depth_limit(oracle):
    b = gen.plummer(3000, seed=4)
    for k in range(1, 6):  # five groups of coincident bodies
        b.x[k * 100:k * 100 + k + 1] = b.x[k * 100]
        b.y[k * 100:k * 100 + k + 1] = b.y[k * 100]
        b.z[k * 100:k * 100 + k + 1] = b.z[k * 100]
    st = oracle.State.from_bodies(b)
    for max_depth in (21, 5):
        for theta in (0.0, 0.3):
            f, depth, stats, _ = run(b, theta=theta, max_depth=max_depth)
            fo, do, info = oracle.bh_force(st, G=1.0, theta=theta, max_depth=max_depth)
            np.testing.assert_array_equal(depth, do)
            assert stats["buckets"] == info["buckets"] > 0
            if theta == 0.0:
                np.testing.assert_array_equal(f, fo)
            else:
                assert rel_diff(f, fo).max() <= 1e-11


@pytest.mark.parametrize("n", [0, 1, 2, 3])
def test_tiny_collections(oracle, n):
    b = gen.uniform_sphere(n, seed=5)
    st = oracle.State.from_bodies(b)
    with NBodyGPU(G=1.0, precision="fp64", collisions=False) as g:
        g.load(b)
        g.bh_enable()
        f = g.forces()
        if n:
            np.testing.assert_array_equal(g.bh_depths(), oracle.bh_force(st, G=1.0)[1])
    np.testing.assert_array_equal(f, oracle.bh_force(st, G=1.0)[0])


@pytest.mark.parametrize("radius_k,mixed_beh", [(0.2, False), (1.0, True)])
def test_collision_pairs_match_oracle(oracle, radius_k, mixed_beh):
    b = gen.two_plummer(20000, seed=6, radius_k=radius_k)
    if mixed_beh:
        b.beh[:] = np.random.default_rng(7).integers(0, 4, b.n)
        b.flags[::13] = 0
    st = oracle.State.from_bodies(b)
    _, _, _, pairs = run(b, collisions=True)
    exp = oracle.detect(st)
    assert len(exp) > 50
    np.testing.assert_array_equal(pairs, exp)


def test_step_with_collisions_matches_oracle(oracle):
    b = gen.two_plummer(20000, seed=8, radius_k=0.2)
    st = oracle.State.from_bodies(b)
    f, _, _ = oracle.bh_force(st, G=1.0)
    pairs = oracle.detect(st)
    oracle.resolve(st, pairs)
    oracle.update(st, f, 1e-3, 0.9)
    with NBodyGPU(G=1.0, precision="fp64", collisions=True) as g:
        g.load(b)
        g.bh_enable()
        s = g.step(1e-3, 0.9, 1)
        out = g.store()
    assert s["pairs_last"] == len(pairs) > 20
    np.testing.assert_allclose(out.vel(), st.vel(), rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(out.pos(), st.pos(), rtol=1e-13, atol=1e-15)


def test_rejected_configurations():
    b = gen.plummer(1000, seed=9)
    with NBodyGPU(G=1.0, eps=0.01, overlap_cutoff=False, collisions=False) as g:  # FP32 path
        g.load(b)
        with pytest.raises(NBodyError) as e:
            g.bh_enable()
        assert e.value.code == L.NB_E_UNSUPPORTED
    with NBodyGPU(G=1.0, precision="fp64") as g:
        g.load(b)
        for bad in (dict(theta=-1.0), dict(max_depth=0), dict(max_depth=22),
                    dict(softening=float("nan")), dict(term=3), dict(term=-1)):
            with pytest.raises(NBodyError) as e:
                g.bh_enable(**bad)
            assert e.value.code == L.NB_E_INVALID


def test_tree_after_a_device_cycle_matches_a_fresh_context():
    b = gen.plummer(20000, seed=10)
    with NBodyGPU(G=1.0, precision="fp64", collisions=False) as g:
        g.load(b)
        g.bh_enable()
        g.forces()
        g.remove_bodies(1500)
        g.add_bodies(gen.plummer(7000, seed=11))
        g.cycle(1.0)
        state = g.store()
        f = g.forces()
        depth = g.bh_depths()
    with NBodyGPU(G=1.0, precision="fp64", collisions=False) as h:
        h.load(state)
        h.bh_enable()
        np.testing.assert_array_equal(h.forces(), f)
        np.testing.assert_array_equal(h.bh_depths(), depth)


# ---------------------------------------------------------------------------------------------
# Fast mode: one walk per 64-body group (opening what any of them would open), FP32 terms
# ---------------------------------------------------------------------------------------------
def run_modes(b, theta=0.3, max_depth=21):
    out = {}
    with NBodyGPU(G=1.0, precision="fp64", collisions=False) as g:
        g.load(b)
        for mode in ("fast", "exact"):
            g.bh_enable(theta=theta, max_depth=max_depth, mode=mode)
            out[mode] = (g.forces(), g.bh_stats())
    return out


def glob(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("kind,n", [("plummer", 65536), ("two_plummer", 30000),
                                    ("uniform_sphere", 30000)])
def test_fast_mode_against_the_fp64_direct_sum(oracle, kind, n):
    """Error against the FP64 softened direct sum (octree.go:314-325 over every body) not above
    exact mode's: the group test opens at least what each body's own test opens.  On the
    Plummer sphere, within 2.5e-4.  Removed and fragmenting bodies feel nothing."""
    b = mixed(getattr(gen, kind)(n, seed=7), 7)
    st = oracle.State.from_bodies(b)
    r = run_modes(b)
    idx = np.arange(0, b.n, 31)
    ref = oracle.force_soft(st, 1.0, 1e-4, idx=idx, nthreads=16)
    act = ((b.flags & L.NB_F_EXISTS) != 0) & ((b.flags & L.NB_F_FRAGMENTING) == 0)
    sel = act[idx]
    e_fast = glob(r["fast"][0][idx][sel], ref[sel])
    e_exact = glob(r["exact"][0][idx][sel], ref[sel])
    print("%s N=%d: fast %.3e, exact %.3e, terms fast/exact %.2f" % (
        kind, n, e_fast, e_exact, r["fast"][1]["interactions"] / r["exact"][1]["interactions"]),
        flush=True)
    if kind == "plummer":
        assert e_fast <= 2.5e-4
    assert e_fast <= e_exact * 1.05 + 2e-6
    assert np.all(r["fast"][0][~act] == 0)
    assert r["fast"][1]["overflow"] == 0


def test_fast_mode_theta_zero_is_an_fp32_direct_sum(oracle):
    b = mixed(gen.plummer(20000, seed=8), 8)
    st = oracle.State.from_bodies(b)
    r = run_modes(b, theta=0.0)
    ref = oracle.force_soft(st, 1.0, 1e-4, nthreads=16)
    act = ((b.flags & L.NB_F_EXISTS) != 0) & ((b.flags & L.NB_F_FRAGMENTING) == 0)
    assert glob(r["fast"][0][act], ref[act]) <= 1e-5
    # every other active body is one term of every active body
    na = int(act.sum())
    assert r["fast"][1]["interactions"] == na * na


def test_fast_mode_suns_and_buckets(oracle):
    """Suns keep the exact direct sums (bit-identical to exact mode); coincident groups under a
    low depth limit are buckets whose bodies are terms one by one."""
    b = gen.plummer(6000, seed=9)
    for k in range(1, 6):
        b.x[k * 100:k * 100 + k + 1] = b.x[k * 100]
        b.y[k * 100:k * 100 + k + 1] = b.y[k * 100]
        b.z[k * 100:k * 100 + k + 1] = b.z[k * 100]
    b.flags[[1, 99, 5999]] |= L.NB_F_SUN
    b.m[[1, 99, 5999]] *= 1e3
    st = oracle.State.from_bodies(b)
    ref = oracle.force_soft(st, 1.0, 1e-4, nthreads=16)
    for max_depth in (21, 5):
        r = run_modes(b, max_depth=max_depth)
        np.testing.assert_array_equal(r["fast"][0][[1, 99, 5999]], r["exact"][0][[1, 99, 5999]])
        assert r["fast"][1]["buckets"] > 0
        assert glob(r["fast"][0], ref) <= 2.5e-4
