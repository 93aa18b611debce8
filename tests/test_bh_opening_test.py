"""The exact octree walk's opening test decided from squares (nb_bh.hip, bh_quick_test).

The reference opens or accepts a node by RN(2 half / RN(sqrt(distSq))) < theta
(cmd/body/octree.go:299-306).  The GPU walk first compares a = RN((2h)^2) against
b = RN(RN(theta^2) distSq) with a relative margin of 2^-40 and only evaluates the reference's
expression when neither a < b (1 - 2^-40) nor a > b (1 + 2^-40) holds, or when b lies outside
[2^-1000, 2^1000] (the host passes NaN for theta^2 when that is not a normal number).  These
tests restate that decision in numpy and check it against the reference's expression on
inputs built to sit within a few ulps of the boundary, at extreme magnitudes and at special
values: a decided case must never disagree.  The same inputs with no margin must disagree
somewhere, which shows they reach the boundary closely enough to catch a margin that is too
small.
"""
import numpy as np

LO, HI = 1.0 - 2.0**-40, 1.0 + 2.0**-40


def theta2_arg(theta):
    t2 = theta * theta
    return t2 if 2.0**-1000 <= t2 <= 2.0**1000 else np.nan


def quick(half, dist_sq, theta2, lo=LO, hi=HI):
    with np.errstate(over="ignore", invalid="ignore", under="ignore"):
        h2 = 2 * half
        a = h2 * h2
        b = theta2 * dist_sq
        normal = (b >= 2.0**-1000) & (b <= 2.0**1000)
        return normal & (a < b * lo), normal & (a > b * hi)


def reference(half, dist_sq, theta):
    with np.errstate(over="ignore", invalid="ignore", divide="ignore", under="ignore"):
        return (2 * half) / np.sqrt(dist_sq) < theta


def near_boundary(rng, theta, scale, n):
    dist_sq = np.exp(rng.uniform(-3, 3, n)) * scale**2
    dist = np.sqrt(dist_sq)
    k = rng.integers(-24, 25, n).astype(float)
    half = theta * dist / 2 * (1 + k * 2.0**-52)
    far = rng.random(n) < 0.2
    half[far] *= np.exp(rng.uniform(-1, 1, far.sum()))
    return half, dist_sq


def check(half, dist_sq, theta, lo=LO, hi=HI):
    yes, no = quick(half, dist_sq, theta2_arg(theta), lo, hi)
    ref = reference(half, dist_sq, theta)
    return int(np.sum(yes & ~ref) + np.sum(no & ref)), int(np.sum(yes | no))


def test_decided_cases_agree_near_the_boundary():
    rng = np.random.default_rng(11)
    for theta in (0.3, 0.5, 0.7, 1.0, 0.1, 1.7, 0.125):
        for scale in (1e-6, 1.0, 1e11, 1e-150, 1e150):
            half, dist_sq = near_boundary(rng, theta, scale, 400_000)
            bad, decided = check(half, dist_sq, theta)
            assert bad == 0, (theta, scale)
            assert decided > 0


def test_without_the_margin_these_inputs_disagree():
    rng = np.random.default_rng(12)
    bad = 0
    for theta in (0.3, 0.7, 1.7):
        half, dist_sq = near_boundary(rng, theta, 1.0, 400_000)
        bad += check(half, dist_sq, theta, lo=1.0, hi=1.0)[0]
    assert bad > 0


def test_extreme_and_special_values_are_never_misdecided():
    rng = np.random.default_rng(13)
    exps = np.concatenate([rng.uniform(-1074, 1023, 200_000), [-1074, -1022, -1000, 0, 1000, 1023]])
    dist_sq = np.exp2(exps)
    half = np.exp2(rng.uniform(-1074, 1023, dist_sq.size))
    specials = np.array([0.0, np.inf, np.nan, 5e-324, 2.0**-1022, 2.0**1023])
    hs, ds = np.meshgrid(specials, specials)
    half = np.concatenate([half, hs.ravel(), rng.choice(specials, 1000)])
    dist_sq = np.concatenate([dist_sq, ds.ravel(), np.exp2(rng.uniform(-1074, 1023, 1000))])
    for theta in (0.3, 1.0, 0.0, 1e-200, 1e200):
        bad, _ = check(half, dist_sq, theta)
        assert bad == 0, theta


def test_theta_zero_leaves_every_node_to_the_reference_expression():
    # theta^2 = 0 is not normal: the host passes NaN and nothing is decided from squares
    yes, no = quick(np.array([1.0, 0.0]), np.array([1.0, 1.0]), theta2_arg(0.0))
    assert not yes.any() and not no.any()
