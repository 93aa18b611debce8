"""Pins of the Barnes-Hut oracle (oracle/nb_oracle_bh.c, cmd/body/octree.go with Build wired in).

theta = 0 visits every leaf, so it must equal the softened direct sum (long double brute force);
hand-worked trees fix the topology (the reference's centre arithmetic, capacity-1 leaves,
re-insertion on subdivision); a huge theta accepts the root, so every body feels the monopole of
the whole set (closed form); a distant cluster is seen as the monopole of the node holding it;
suns sum directly; bodies that do not exist or fragment are neither targets nor sources.
"""
import numpy as np
import pytest

from nbodygo_amd import gen

S = 1e-4  # Softening (octree.go:14)


def state(oracle, pos, m, flags=None):
    pos = np.asarray(pos, float)
    n = len(pos)
    z = np.zeros(n)
    return oracle.State(pos[:, 0], pos[:, 1], pos[:, 2], z, z, z, np.asarray(m, float), z,
                        flags=flags)


def soft_ld(pos, m, G, targets=None, exclude=None):
    """Long-double softened direct sum on each target from every other source."""
    pos = np.asarray(pos, np.longdouble)
    m = np.asarray(m, np.longdouble)
    n = len(pos)
    out = np.zeros((n, 3), np.longdouble)
    for i in (range(n) if targets is None else targets):
        d = pos - pos[i]
        r2 = (d * d).sum(1) + np.longdouble(S) * np.longdouble(S)
        w = np.longdouble(G) * m[i] * m / (r2 * np.sqrt(r2))
        w[i] = 0
        if exclude is not None:
            w[exclude] = 0
        out[i] = (w[:, None] * d).sum(0)
    return out.astype(float)


def test_theta_zero_is_the_softened_direct_sum(oracle):
    b = gen.plummer(3000, seed=1)
    st = oracle.State.from_bodies(b)
    f, depth, info = oracle.bh_force(st, G=1.0, theta=0.0)
    ref = soft_ld(b.pos(), b.m, 1.0)
    rel = np.linalg.norm(f - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert rel.max() < 1e-12
    assert info["buckets"] == 0 and info["interactions"] == 3000 * 2999
    assert (depth >= 1).all() and info["max_depth_seen"] == depth.max()


def test_hand_worked_three_body_tree(oracle):
    # box x in [0, 1]: centre (0.5, 0, 0), half 0.525.  Bodies 0 and 2 share octant 6 down to
    # the node centred at x = 0.040625 (depth 3), where they part: leaves at depth 4.
    pos = [(0.0, 0, 0), (1.0, 0, 0), (0.1, 0, 0)]
    st = state(oracle, pos, [1.0, 2.0, 3.0])
    f, depth, info = oracle.bh_force(st, G=1.0, theta=0.0)
    assert list(depth) == [4, 1, 4]
    assert info["internal"] == 4 and info["nodes"] == 1 + 8 * 4
    assert info["root_half"] == 0.525
    assert info["root_mass"] == 6.0
    np.testing.assert_allclose(f, soft_ld(pos, [1, 2, 3], 1.0), rtol=1e-14)


def test_two_bodies_and_the_leaf_centre_of_mass(oracle):
    # a fresh leaf's centre of mass is (x*m)/m (updateMass from zero), used as the source point
    pos = [(0.3, -0.7, 0.11), (-0.2, 0.5, 0.9)]
    m = [0.37, 1.9]
    st = state(oracle, pos, m)
    f, depth, info = oracle.bh_force(st, G=2.5, theta=0.3)
    src = [(np.array(p) * mm) / mm for p, mm in zip(pos, m)]
    for i, j in ((0, 1), (1, 0)):
        d = src[j] - np.array(pos[i])
        d2 = d @ d + S * S
        exp = 2.5 * m[i] * m[j] / d2 * d / np.sqrt(d2)
        np.testing.assert_allclose(f[i], exp, rtol=1e-15)
    assert list(depth) == [1, 1] and info["nodes"] == 9


def test_huge_theta_accepts_the_root(oracle):
    b = gen.uniform_sphere(500, seed=2)
    st = oracle.State.from_bodies(b)
    f, _, info = oracle.bh_force(st, G=1.0, theta=1e30)
    M = np.sum(b.m, dtype=np.longdouble)
    cm = (b.pos().astype(np.longdouble) * b.m[:, None]).sum(0) / M
    assert abs(info["root_mass"] - float(M)) <= 1e-14 * float(M)
    np.testing.assert_allclose(info["root_cm"], cm.astype(float), rtol=0, atol=1e-13)
    d = cm - b.pos().astype(np.longdouble)
    d2 = (d * d).sum(1) + np.longdouble(S) ** 2
    exp = (b.m[:, None] * M * d / (d2 * np.sqrt(d2))[:, None]).astype(float)
    np.testing.assert_allclose(f, exp, rtol=1e-9, atol=1e-12 * np.abs(exp).max())
    assert info["interactions"] == 500


def test_distant_cluster_is_its_monopole(oracle):
    rng = np.random.default_rng(3)
    # one node at depth 2 holds the whole cluster (centre (895.6, 232.4, 232.4), half 131.6),
    # and 2 * 131.6 / 1020 < 0.3 accepts it for the probe at the origin
    cl = rng.normal(0, 1, (200, 3)) + np.array([1000.0, 200.0, 200.0])
    pos = np.vstack([[[0.0, 0.0, 0.0]], cl])
    m = np.concatenate([[1.0], rng.uniform(0.5, 2.0, 200)])
    st = state(oracle, pos, m)
    f, _, _ = oracle.bh_force(st, G=1.0, theta=0.3)
    M = m[1:].astype(np.longdouble).sum()
    cm = (cl.astype(np.longdouble) * m[1:, None]).sum(0) / M
    d = cm - pos[0]
    d2 = (d * d).sum() + np.longdouble(S) ** 2
    exp = (np.longdouble(m[0]) * M * d / (d2 * np.sqrt(d2))).astype(float)
    np.testing.assert_allclose(f[0], exp, rtol=1e-12)
    # while the direct sum differs from the monopole by the cluster's quadrupole (~1e-6 here)
    direct = soft_ld(pos, m, 1.0, targets=[0])[0]
    assert 1e-9 < np.linalg.norm(direct - exp) / np.linalg.norm(exp) < 1e-4


def test_suns_sum_directly_and_inactive_bodies_are_skipped(oracle):
    b = gen.plummer(1500, seed=4)
    b.flags[7] |= 4     # a sun
    b.flags[::50] = 0   # not existing
    b.flags[3::61] |= 2  # fragmenting
    b.flags[7] |= 1
    st = oracle.State.from_bodies(b)
    f, depth, _ = oracle.bh_force(st, G=1.0, theta=0.3)
    act = ((b.flags & 1) != 0) & ((b.flags & 2) == 0)
    assert np.all(f[~act] == 0) and np.all(depth[~act] == -1)
    exp = soft_ld(b.pos(), b.m, 1.0, targets=[7], exclude=~act)[7]
    np.testing.assert_allclose(f[7], exp, rtol=1e-13)
    # the sun is still a source in everyone else's tree: theta = 0 equals the direct sum
    f0, _, _ = oracle.bh_force(st, G=1.0, theta=0.0)
    ref = soft_ld(b.pos(), b.m, 1.0, targets=np.nonzero(act)[0][:50], exclude=~act)
    for i in np.nonzero(act)[0][:50]:
        np.testing.assert_allclose(f0[i], ref[i], rtol=1e-11)


def test_coincident_bodies_share_a_bucket_at_the_depth_limit(oracle):
    pos = [(0.0, 0, 0), (0.5, 0.5, 0.5), (0.5, 0.5, 0.5), (0.5, 0.5, 0.5), (1.0, 0, 1.0)]
    m = [1.0, 2.0, 3.0, 4.0, 5.0]
    st = state(oracle, pos, m)
    f, depth, info = oracle.bh_force(st, G=1.0, theta=0.0, max_depth=21)
    assert info["buckets"] == 1 and list(depth) == [1, 21, 21, 21, 1]
    np.testing.assert_allclose(f, soft_ld(pos, m, 1.0), rtol=1e-13, atol=1e-300)


@pytest.mark.parametrize("theta,bar", [(0.3, 2e-3), (0.1, 2e-4)])
def test_error_against_the_direct_sum_falls_with_theta(oracle, theta, bar):
    b = gen.plummer(4000, seed=5)
    st = oracle.State.from_bodies(b)
    f, _, _ = oracle.bh_force(st, G=1.0, theta=theta)
    ref = oracle.force_soft(st, 1.0, S)
    assert np.linalg.norm(f - ref) / np.linalg.norm(ref) < bar
