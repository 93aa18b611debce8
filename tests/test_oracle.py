"""Pins the CPU FP64 oracle against closed forms, golden vectors and brute force (CPU only).

Each check is chosen so that a plausible slip in the oracle (a dropped term, a wrong sign, a
swapped operand, a wrong index) fails it."""
from decimal import Decimal as D

import numpy as np
import pytest

from nbodygo_amd import gen

G_REF = 6.673e-11


def two_body_state(oracle, golden):
    w = golden["wp_compute"]
    p = np.array(w["pos"], float)
    return oracle.State(p[:, 0], p[:, 1], p[:, 2], [0, 0], [0, 0], [0, 0], w["m"], w["r"])


def test_wp_compute_force_closed_form(oracle, golden):
    """TestWpCompute geometry (workpool_test.go:41-56): force on both bodies."""
    st = two_body_state(oracle, golden)
    f = oracle.force_ref(st, G_REF)
    w = golden["wp_compute"]
    for c in range(3):
        assert f[0, c] == pytest.approx(float(D(w["f0"][c])), rel=4e-16)
        assert f[1, c] == pytest.approx(float(D(w["f1"][c])), rel=4e-16)
    assert f[0, 0] > 0 and f[1, 0] < 0  # attraction: body 0 is pulled towards +x


def test_wp_compute_update_closed_form(oracle, golden):
    """Update(1, 1) after the force (workpool_test.go:51-52): V = f, X = X0 + V."""
    st = two_body_state(oracle, golden)
    f = oracle.force_ref(st, G_REF)
    oracle.update(st, f, 1.0, 1.0)
    w = golden["wp_compute"]
    assert st.vx[0] == pytest.approx(float(D(w["v0_after_update"][0])), rel=4e-16)
    assert st.x[0] == float(D(w["x0_after_update"][0]))
    assert st.vx[1] == -st.vx[0]


def test_newton_third_law(oracle):
    b = gen.uniform_sphere(300, seed=3, radius_k=0.02)
    st = oracle.State.from_bodies(b)
    for f in (oracle.force_ref(st, G_REF), oracle.force_soft(st, 1.0, 30.0)):
        scale = np.abs(f).sum()
        assert np.abs(f.sum(axis=0)).max() < 1e-13 * scale


def test_soft_force_against_long_double(oracle):
    b = gen.plummer(700, seed=5)
    st = oracle.State.from_bodies(b)
    f = oracle.force_soft(st, 1.0, 0.01)
    fl = oracle.force_soft(st, 1.0, 0.01, long_double=True)
    err = np.linalg.norm(f - fl, axis=1) / np.linalg.norm(fl, axis=1)
    assert err.max() < 1e-12


def test_ref_force_against_long_double_with_cutoff(oracle):
    """Reference semantics: the unsoftened sum with the overlap gate equals the eps = 0 long-
    double sum with the same gate (radii large enough that many pairs are gated out)."""
    b = gen.uniform_sphere(500, seed=11, radius_k=0.3)
    st = oracle.State.from_bodies(b)
    f = oracle.force_ref(st, G_REF)
    fl = oracle.force_soft(st, G_REF, 0.0, cutoff=True, long_double=True)
    err = np.linalg.norm(f - fl, axis=1) / np.maximum(np.linalg.norm(fl, axis=1), 1e-300)
    assert err.max() < 1e-12
    # the gate actually removed pairs in this geometry
    assert len(oracle.detect(st)) > 50


def test_cutoff_removes_exactly_the_overlapping_pairs(oracle):
    """Two bodies at distance exactly Ri + Rj overlap (d <= R1 + R2 is a hit, body.go:200)."""
    st = oracle.State([0.0, 3.0], [0.0, 4.0], [0.0, 0.0], [0, 0], [0, 0], [0, 0], [1.0, 1.0],
                      [2.5, 2.5])
    assert np.all(oracle.force_ref(st, 1.0) == 0.0)
    st2 = oracle.State([0.0, 3.0], [0.0, 4.0], [0.0, 0.0], [0, 0], [0, 0], [0, 0], [1.0, 1.0],
                       [2.5, 2.4999999])
    f = oracle.force_ref(st2, 1.0)
    assert f[0, 0] == pytest.approx(1.0 / 25 * 3 / 5, rel=1e-15)
    assert f[0, 1] == pytest.approx(1.0 / 25 * 4 / 5, rel=1e-15)


def test_symmetric_lattice_centre_force_vanishes(oracle):
    g = np.array([-1.0, 0.0, 1.0])
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    x, y, z = X.ravel(), Y.ravel(), Z.ravel()
    n = x.size
    st = oracle.State(x, y, z, np.zeros(n), np.zeros(n), np.zeros(n), np.ones(n),
                      np.full(n, 0.1))
    f = oracle.force_ref(st, 1.0)
    centre = 13
    assert np.abs(f[centre]).max() < 1e-15
    # a corner is pulled towards the centre along the diagonal
    assert np.all(np.sign(f[0]) == 1)


def test_plummer_field_matches_analytic(oracle):
    """Mean radial acceleration in shells against a(r) = -G M r / (r^2 + a^2)^(3/2)."""
    n = 100000
    b = gen.plummer(n, seed=9)
    st = oracle.State.from_bodies(b)
    rng = np.random.default_rng(1)
    idx = rng.choice(n, 3000, replace=False)
    f = oracle.force_soft(st, 1.0, 1e-3, idx=idx)
    a = f / b.m[idx, None]
    p = b.pos()[idx]
    r = np.linalg.norm(p, axis=1)
    ar = np.sum(a * p, axis=1) / r
    ana = -r / (r * r + 1.0) ** 1.5
    for lo, hi in ((0.3, 0.7), (0.7, 1.2), (1.2, 2.0)):
        sel = (r >= lo) & (r < hi)
        assert abs(ar[sel].mean() / ana[sel].mean() - 1.0) < 0.03


def brute_pairs(b):
    """Detection by the formula of Collided (body.go:196-202) evaluated with numpy."""
    act = ((b.flags & 1) != 0) & ((b.flags & 2) == 0)
    cc = (b.beh == 2) | (b.beh == 3)
    out = []
    for i in range(b.n):
        if not (act[i] and cc[i]):
            continue
        dx = b.x[i + 1:] - b.x[i]
        dy = b.y[i + 1:] - b.y[i]
        dz = b.z[i + 1:] - b.z[i]
        d = np.sqrt(dx * dx + dy * dy + dz * dz)
        hit = ~(d > b.r[i] + b.r[i + 1:]) & act[i + 1:] & cc[i + 1:]
        out += [(i, i + 1 + j) for j in np.nonzero(hit)[0]]
    return np.array(out, np.int32).reshape(-1, 2)


def test_detect_matches_brute_force_with_edge_cases(oracle):
    b = gen.uniform_sphere(600, seed=4, radius_k=0.12)
    # boundary: d == R1 + R2 exactly (3-4-5 triangle), d == 0, inactive and None bodies
    b.x[:4] = [0.0, 3.0, 0.0, 0.0]
    b.y[:4] = [0.0, 4.0, 0.0, 0.0]
    b.z[:4] = [5000.0, 5000.0, 6000.0, 6000.0]
    b.r[:4] = [2.5, 2.5, 1.0, 1.0]
    b.flags[10] = 0           # not existing
    b.flags[11] |= 2          # fragmenting
    b.beh[12] = 0             # None: no collision events
    b.beh[13] = 3             # Fragment still collides
    st = oracle.State.from_bodies(b)
    got = oracle.detect(st)
    exp = brute_pairs(b)
    assert got.shape[0] > 20
    np.testing.assert_array_equal(got, exp)
    pairs = {tuple(p) for p in got}
    assert (0, 1) in pairs and (2, 3) in pairs


@pytest.mark.parametrize("scene", ["edge_cases", "two_spheres", "mixed_radii", "cell_faces"])
def test_detect_grid_matches_brute_force(oracle, scene):
    """The cell-list detector gives the brute-force detector's canonical pair list: boundary
    geometries (d == R1 + R2, coincident), inactive / None / Fragment bodies, two clusters far
    apart (cells of very different occupancy), radii of widely different size, and pairs placed
    across cell faces at exactly the cell width."""
    if scene == "edge_cases":
        b = gen.uniform_sphere(600, seed=4, radius_k=0.12)
        b.x[:4] = [0.0, 3.0, 0.0, 0.0]
        b.y[:4] = [0.0, 4.0, 0.0, 0.0]
        b.z[:4] = [5000.0, 5000.0, 6000.0, 6000.0]
        b.r[:4] = [2.5, 2.5, 1.0, 1.0]
        b.flags[10] = 0
        b.flags[11] |= 2
        b.beh[12] = 0
        b.beh[13] = 3
    elif scene == "two_spheres":
        b = gen.two_plummer(1500, seed=5, radius_k=0.6)
    elif scene == "mixed_radii":
        b = gen.uniform_sphere(2000, seed=6, radius_k=0.05)
        b.r[::97] *= 20.0
    else:
        rng = np.random.default_rng(7)
        n = 1200
        pos = rng.uniform(0.0, 8.0, (n, 3)).round(0)  # a lattice of unit spacing
        pos[n // 2:] += [1.0, 0.0, 0.0]               # duplicated one cell over: d == 1 == 2 r
        from nbodygo_amd import Bodies
        b = Bodies(pos[:, 0], pos[:, 1], pos[:, 2], *np.zeros((3, n)), np.ones(n),
                   np.full(n, 0.5))
    st = oracle.State.from_bodies(b)
    exp = oracle.detect(st)
    assert len(exp) > 20
    for threads in (1, 5):
        np.testing.assert_array_equal(oracle.detect_grid(st, nthreads=threads), exp)


def golden_state(oracle, c):
    x1, x2, v1, v2 = c["x1"], c["x2"], c["v1"], c["v2"]
    st = oracle.State([x1[0], x2[0]], [x1[1], x2[1]], [x1[2], x2[2]], [v1[0], v2[0]],
                      [v1[1], v2[1]], [v1[2], v2[2]], [c["m1"], c["m2"]], [c["r1"], c["r2"]],
                      rest=[c["R"], c["R"]])
    return st


@pytest.mark.parametrize("name", ["head_on", "one_d"])
def test_elastic_closed_forms(oracle, golden, name):
    c = golden[name]
    st = golden_state(oracle, c)
    oracle.resolve(st, [[0, 1]])
    for k, ax in enumerate(("vx", "vy", "vz")):
        assert getattr(st, ax)[0] == pytest.approx(float(D(c["v1_after"][k])), abs=1e-14)
        assert getattr(st, ax)[1] == pytest.approx(float(D(c["v2_after"][k])), abs=1e-14)
    assert st.collided.tolist() == [1, 1]


def test_elastic_oblique_contact_normal(oracle, golden):
    """coll3d equals the impulse along the line of centres at first contact."""
    for c in golden["oblique"]:
        st = golden_state(oracle, c)
        oracle.resolve(st, [[0, 1]])
        got = np.array([[st.vx[0], st.vy[0], st.vz[0]], [st.vx[1], st.vy[1], st.vz[1]]])
        exp = np.array([[float(D(q)) for q in c["v1_after"]],
                        [float(D(q)) for q in c["v2_after"]]])
        scale = np.abs(np.array([c["v1"], c["v2"]])).max()
        assert np.abs(got - exp).max() < 1e-10 * scale, c


def test_elastic_invariants(oracle):
    rng = np.random.default_rng(2)
    n_hit = 0
    for _ in range(300):
        x = rng.uniform(-1, 1, (2, 3))
        v = rng.uniform(-2, 2, (2, 3))
        m = rng.uniform(0.1, 4, 2)
        R = rng.choice([1.0, 0.5])
        st = oracle.State(x[:, 0], x[:, 1], x[:, 2], v[:, 0], v[:, 1], v[:, 2], m, [1.0, 1.0],
                          rest=[R, R])
        ok, out = oracle.elastic(st, 0, 1)
        oracle.resolve(st, [[0, 1]])
        v2 = st.vel()
        p0 = m[0] * v[0] + m[1] * v[1]
        p1 = m[0] * v2[0] + m[1] * v2[1]
        assert np.abs(p1 - p0).max() < 1e-12 * (np.abs(m[:, None] * v).sum())
        if ok:
            n_hit += 1
            ke0 = 0.5 * np.sum(m * np.sum(v * v, axis=1))
            ke1 = 0.5 * np.sum(m * np.sum(v2 * v2, axis=1))
            if R == 1.0:
                assert ke1 == pytest.approx(ke0, rel=1e-12)
            else:
                assert ke1 < ke0
            # swap symmetry: resolving (1, 0) on the original state gives the same velocities
            st2 = oracle.State(x[:, 0], x[:, 1], x[:, 2], v[:, 0], v[:, 1], v[:, 2], m,
                               [1.0, 1.0], rest=[R, R])
            oracle.resolve(st2, [[1, 0]])
            assert np.abs(st2.vel() - v2).max() < 1e-9 * np.abs(v).max()
        else:
            assert np.array_equal(v2, v)
    assert n_hit > 50


def test_elastic_no_ops(oracle):
    # equal velocities: v == 0 exit (collisioncalc.go:89-92)
    st = oracle.State([0, 1], [0, 0], [0, 0], [1, 1], [2, 2], [0, 0], [1, 2], [1, 1])
    assert oracle.elastic(st, 0, 1)[0] is False
    # separating: thetav > pi/2 exit (collisioncalc.go:137-140)
    st = oracle.State([0, 1], [0, 0], [0, 0], [-1, 1], [0, 0], [0, 0], [1, 1], [1, 1])
    assert oracle.elastic(st, 0, 1)[0] is False
    st.vx[:] = [1, -1]
    assert oracle.elastic(st, 0, 1)[0] is True


def test_collide_coincident_bodies_go_nan_and_are_removed(oracle):
    """TestCollide (body_collection_test.go:319-344): coincident bodies are marked collided;
    acos(0/0) makes their velocities NaN and the next Update removes both (body.go:134-137)."""
    st = oracle.State([500, 500], [500, 500], [500, 500], [1, 2], [1, 2], [1, 2], [1, 2], [1, 2])
    oracle.resolve(st, [[0, 1]])
    assert st.collided.tolist() == [1, 1]
    assert np.isnan(st.vx).all()
    s = oracle.update(st, np.zeros((2, 3)), 1.0, 1.0)
    assert s["removed_total"] == 2
    assert (st.flags & 1).tolist() == [0, 0]


def test_update_rules(oracle):
    n = 4
    st = oracle.State(np.zeros(n), np.zeros(n), np.zeros(n), [1.0, 2.0, 3.0, 4.0], np.zeros(n),
                      np.zeros(n), [2.0, 2.0, 0.0, 1.0], np.zeros(n))
    st.collided[1] = 1
    st.flags[3] = 0
    f = np.array([[4.0, 0, 0], [4.0, 0, 0], [0.0, 0, 0], [4.0, 0, 0]])
    s = oracle.update(st, f, 0.5, 0.25)
    assert st.vx[0] == 1.0 + 0.5 * 4.0 / 2.0 and st.x[0] == 0.5 * st.vx[0]   # kick then drift
    assert st.vx[1] == 2.0 and st.x[1] == 1.0                                # collided: no kick
    assert np.isnan(st.vx[2]) and (st.flags[2] & 1) == 0                     # zero mass: 0/0
    assert st.vx[3] == 4.0 and st.x[3] == 0.0                                # not existing
    assert st.rest[0] == 0.25 and st.rest[3] == 1.0
    assert st.collided.tolist() == [0, 0, 0, 0]
    assert s["removed_total"] == 1


def test_free_drift(oracle):
    st = oracle.State([1.0], [2.0], [3.0], [0.5], [-0.25], [0.125], [1.0], [0.0])
    x = 1.0
    for _ in range(64):
        oracle.step(st, 1.0, 0.0, True, True, 0.125, 1.0)
        x += 0.125 * 0.5
    assert st.x[0] == x == 1.0 + 64 * 0.125 * 0.5


def test_step_composes_the_phases(oracle):
    """or_step == force, detect, resolve, update called one by one (bit for bit)."""
    b = gen.uniform_sphere(400, seed=8, radius_k=0.1)
    s1, s2 = oracle.State.from_bodies(b), oracle.State.from_bodies(b)
    for _ in range(3):
        out = oracle.step(s1, G_REF, 0.0, True, True, 1e-3, 0.9)
        f = oracle.force_ref(s2, G_REF)
        pairs = oracle.detect(s2)
        np.testing.assert_array_equal(out["pairs"], pairs)
        oracle.resolve(s2, pairs)
        oracle.update(s2, f, 1e-3, 0.9)
        for k in ("x", "y", "z", "vx", "vy", "vz"):
            np.testing.assert_array_equal(getattr(s1, k), getattr(s2, k))
    assert out["pairs_last"] > 0


def test_energy_two_body(oracle):
    st = oracle.State([0.0, 3.0], [0.0, 4.0], [0.0, 0.0], [1.0, 0.0], [0.0, 2.0], [0.0, 0.0],
                      [2.0, 3.0], [0.0, 0.0])
    e = oracle.energy(st, 1.5, 0.0)
    assert e["kinetic"] == pytest.approx(0.5 * 2 * 1 + 0.5 * 3 * 4)
    assert e["potential"] == pytest.approx(-1.5 * 6 / 5)
    np.testing.assert_allclose(e["momentum"], [2.0, 6.0, 0.0])
    e2 = oracle.energy(st, 1.5, 4.0)
    assert e2["potential"] == pytest.approx(-1.5 * 6 / np.sqrt(25 + 16))


def test_circular_orbit_energy_bounded(oracle):
    """Symplectic Euler on a two-body circular orbit keeps the energy bounded over many orbits."""
    G, M = 1.0, 1.0
    v = np.sqrt(0.5)  # m v^2 / 0.5 = G m m / 1^2
    st = oracle.State([-0.5, 0.5], [0, 0], [0, 0], [0, 0], [-v, v], [0, 0], [M, M], [0.0, 0.0])
    e0 = oracle.energy(st, G, 0.0)
    E0 = e0["kinetic"] + e0["potential"]
    for _ in range(20):
        oracle.step(st, G, 0.0, True, False, 1e-3, 1.0, nsteps=500)
        e = oracle.energy(st, G, 0.0)
        assert abs(e["kinetic"] + e["potential"] - E0) < 1e-3 * abs(E0)


def test_workpool_matches_force_and_counts_hits(oracle):
    """The reference's worker-pool structure computes the same forces; its detection loop hits
    every ordered overlapping pair plus every body's self-pair (body.go:172-177)."""
    b = gen.uniform_sphere(500, seed=12, radius_k=0.1)
    st = oracle.State.from_bodies(b)
    el, hits, f = oracle.bench_workpool(st, G_REF, 500, nsteps=1, nthreads=4, want_force=True)
    np.testing.assert_array_equal(f, oracle.force_ref(st, G_REF))
    assert hits == 2 * len(oracle.detect(st)) + 500
    assert el > 0


@pytest.mark.parametrize("threads", [1, 7])
def test_bench_cycles_equals_the_serial_cycle(oracle, threads):
    """The whole-cycle worker-pool baseline advances the state exactly as or_step does (reference
    semantics, collisions on), so the timed CPU work is the reference's cycle.  The dense scene
    makes collision pairs in most cycles; a zero-mass body turns NaN and is removed."""
    b = gen.uniform_sphere(600, seed=3, radius=150.0)
    b.m[5] = 0.0
    a = oracle.State.from_bodies(b)
    c = oracle.State.from_bodies(b)
    ref = oracle.step(a, 6.673e-11, 0.0, True, True, 1e-3, 1.0, nsteps=12)
    el, st = oracle.bench_cycles(c, 6.673e-11, 1e-3, 1.0, 12, threads)
    assert el >= 0
    assert ref["pairs_total"] > 50
    for k in ("pairs_total", "pairs_last", "resolved_total", "removed_total"):
        assert st[k] == ref[k], k
    for k in ("x", "y", "z", "vx", "vy", "vz", "rest"):
        np.testing.assert_array_equal(c.a[k], a.a[k])
    np.testing.assert_array_equal(c.flags, a.flags)
    assert not (c.flags[5] & oracle.F_EXISTS)


@pytest.mark.parametrize("threads", [1, 4])
def test_soa_line_matches_the_softened_sum(oracle, threads):
    """The benchmark's optimised CPU line (vectorised, threaded, its own operation order) gives
    the oracle's FP64 softened accelerations to FP64 rounding, on any target slice."""
    b = gen.plummer(1500, seed=9)
    st = oracle.State.from_bodies(b)
    ref = oracle.force_soft(st, 2.0, 0.05, nthreads=4) / b.m[:, None]
    a, _ = oracle.soa_soft(b.x, b.y, b.z, b.m, 2.0, 0.05, nthreads=threads)
    per = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert per.max() < 1e-12
    part, _ = oracle.soa_soft(b.x, b.y, b.z, b.m, 2.0, 0.05, t0=100, t1=777, nthreads=threads)
    np.testing.assert_array_equal(part, a[100:777])
