"""GPU parity: libnbgpu.so (through the C-ABI) against the CPU FP64 oracle on the same seeded
inputs.

Bars:
  FP64 path   force and update bit-identical to the oracle (same IEEE operations in the same
              order); pair sets identical; after a resolved collision velocities agree to 1e-12
              relative (Go/glibc/ocml transcendentals differ in the last ulp).
  FP32 path   force error against FP64: global ||da||/||a|| <= 1e-5 and
              per-body p99 <= 1e-4; pair sets identical (detection is exact FP64).
"""
import numpy as np
import pytest

from nbodygo_amd import NBodyGPU, NBodyError, gen
from nbodygo_amd import _lib as L

pytestmark = pytest.mark.gpu

G_REF = 6.673e-11
NT = 16


def rel_err(got, exp):
    num = np.linalg.norm(got - exp, axis=1)
    den = np.linalg.norm(exp, axis=1)
    glob = np.linalg.norm(got - exp) / np.linalg.norm(exp)
    per = num / np.maximum(den, 1e-300)
    return glob, np.percentile(per, 50), np.percentile(per, 99), per.max()


def c1(n=1000, seed=1, radius_k=0.05):
    """Config 1: uniform sphere R = 1e3, lognormal masses (mu = ln 1e8, sigma = 1),
    radii 0.05 m^(1/3), velocities N(0, 5^2)."""
    return gen.uniform_sphere(n, seed=seed, radius_k=radius_k)


# ---------------------------------------------------------------------------------------------
# FP64 path: bit-identical to the oracle
# ---------------------------------------------------------------------------------------------
def test_fp64_force_bit_exact_reference_mode(oracle):
    b = c1()
    st = oracle.State.from_bodies(b)
    with NBodyGPU(precision="fp64") as g:
        g.load(b)
        f = g.forces()
        pairs = g.pairs()
    np.testing.assert_array_equal(f, oracle.force_ref(st, G_REF, nthreads=NT))
    np.testing.assert_array_equal(pairs, oracle.detect(st))
    assert len(pairs) > 10


@pytest.mark.parametrize("cutoff", [False, True])
def test_fp64_force_bit_exact_softened(oracle, cutoff):
    b = gen.plummer(3000, seed=2, radius=0.02)
    st = oracle.State.from_bodies(b)
    with NBodyGPU(G=1.0, eps=0.01, overlap_cutoff=cutoff, precision="fp64") as g:
        g.load(b)
        f = g.forces()
    np.testing.assert_array_equal(f, oracle.force_soft(st, 1.0, 0.01, cutoff=cutoff,
                                                       nthreads=NT))


def test_fp64_steps_bit_exact_without_collisions(oracle):
    b = c1(radius_k=0.02)
    st = oracle.State.from_bodies(b)
    with NBodyGPU(precision="fp64", collisions=False) as g:
        g.load(b)
        s = g.step(1e-3, 1.0, nsteps=100)
        out = g.store()
    oracle.step(st, G_REF, 0.0, True, False, 1e-3, 1.0, nsteps=100, nthreads=NT)
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        np.testing.assert_array_equal(getattr(out, k), getattr(st, k))
    assert s["interactions"] == 100 * 1000 * 999


def test_fp64_steps_with_collisions_track_the_oracle(oracle):
    """Config 1 (reference mode, elastic collisions on), 100 steps, stepping both from the same
    state each cycle: identical pair sets every step; states equal up to transcendental ulps."""
    b = c1()
    with NBodyGPU(precision="fp64") as g:
        g.load(b)
        total_pairs = 0
        for step in range(100):
            cur = g.store()
            st = oracle.State.from_bodies(cur)
            st.collided[:] = 0
            g.step(1e-3, 0.9)
            gp = g.pairs()
            o = oracle.step(st, G_REF, 0.0, True, True, 1e-3, 0.9, nthreads=NT)
            np.testing.assert_array_equal(gp, o["pairs"])
            total_pairs += len(gp)
            nxt = g.store()
            for k in ("x", "y", "z", "vx", "vy", "vz"):
                a, e = getattr(nxt, k), getattr(st, k)
                np.testing.assert_allclose(a, e, rtol=1e-12, atol=1e-12 * np.abs(e).max())
    assert total_pairs > 100


# Config 1's softened variant: softening "3e4 * scale" read with scale = the runner's time scaling
# (dt = 1e-3, the -m flag, main.go:199), so eps = 30 length units against a mean spacing of about
# 100 in the radius-1e3 sphere; radii 0.05 m^(1/3) as in reference mode.
C1_EPS = 3e4 * 1e-3
C1_RADIUS_K = 0.05


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_config1_softened_tracks_the_oracle(oracle, precision):
    """Config 1 with eps = 30, overlap gate and elastic collisions, 100 cycles, each GPU cycle
    against one oracle cycle from the same state: identical pair sets every cycle; FP64 states
    equal up to coll3d's transcendental ulps, FP32 states within the FP32 force error."""
    b = c1(radius_k=C1_RADIUS_K)
    rtol_v = 1e-12 if precision == "fp64" else 1e-6
    with NBodyGPU(eps=C1_EPS, precision=precision) as g:
        g.load(b)
        total_pairs = 0
        for step in range(100):
            cur = g.store()
            st = oracle.State.from_bodies(cur)
            g.step(1e-3, 1.0)
            o = oracle.step(st, G_REF, C1_EPS, True, True, 1e-3, 1.0, nthreads=NT)
            np.testing.assert_array_equal(g.pairs(), o["pairs"])
            total_pairs += len(o["pairs"])
            nxt = g.store()
            np.testing.assert_allclose(nxt.pos(), st.pos(), rtol=1e-12,
                                       atol=1e-12 * np.abs(st.pos()).max())
            np.testing.assert_allclose(nxt.vel(), st.vel(), rtol=rtol_v,
                                       atol=rtol_v * np.abs(st.vel()).max())
    assert total_pairs > 100


# ---------------------------------------------------------------------------------------------
# FP32 path: error against FP64, exact pair sets
# ---------------------------------------------------------------------------------------------
def test_fp32_plummer_force_error_config2(oracle):
    """Config 2 (N = 65,536 Plummer, eps = 0.01): force error vs the FP64 oracle on 2,048
    sampled targets; the device FP64 sampler agrees with the oracle to 1e-13."""
    n = 65536
    b = gen.plummer(n, seed=20)
    st = oracle.State.from_bodies(b)
    idx = np.random.default_rng(0).choice(n, 2048, replace=False)
    exp = oracle.force_soft(st, 1.0, 0.01, idx=idx, nthreads=NT)
    with NBodyGPU(G=1.0, eps=0.01, overlap_cutoff=False, collisions=False) as g:
        g.load(b)
        f = g.forces()[idx]
        a64 = g.accel_fp64(idx)
    glob, p50, p99, mx = rel_err(f, exp)
    print("config2 FP32 force error: global %.3e p50 %.3e p99 %.3e max %.3e" % (glob, p50, p99,
                                                                                mx))
    assert glob <= 1e-5
    assert p99 <= 1e-4
    g64, _, _, m64 = rel_err(a64 * b.m[idx, None], exp)
    assert m64 < 1e-13


@pytest.mark.parametrize("shape", [(1, 512, 4), (1, 256, 4), (2, 256, 8)])
def test_headline_force_error_at_size(oracle, shape):
    """The headline workload (N = 262,144 Plummer, eps = 0.01) in each launch shape the autotuner
    picks at that size (targets per lane, source tile, source slices): force error vs the FP64
    oracle on 1,024 sampled targets, to the config-2 bars."""
    n = 262144
    b = gen.plummer(n, seed=2024)
    st = oracle.State.from_bodies(b)
    idx = np.random.default_rng(1).choice(n, 1024, replace=False)
    exp = oracle.force_soft(st, 1.0, 0.01, idx=idx, nthreads=NT)
    ipl, tile_j, jsplit = shape
    with NBodyGPU(G=1.0, eps=0.01, overlap_cutoff=False, collisions=False, ipl=ipl,
                  tile_j=tile_j, jsplit=jsplit, block=256, variant=2) as g:
        g.load(b)
        f = g.forces()[idx]
    glob, p50, p99, mx = rel_err(f, exp)
    print("N=262144 shape %s FP32 force error: global %.3e p99 %.3e max %.3e" % (shape, glob,
                                                                                 p99, mx))
    assert glob <= 1e-5
    assert p99 <= 1e-4


def test_fp32_reference_mode_force_error(oracle):
    """Config 1 in reference semantics (eps = 0, overlap gate): the gate decisions are exact
    FP64, so every body's force is within FP32 error of the oracle's."""
    b = c1()
    st = oracle.State.from_bodies(b)
    exp = oracle.force_ref(st, G_REF, nthreads=NT)
    with NBodyGPU() as g:
        g.load(b)
        f = g.forces()
        pairs = g.pairs()
    glob, p50, p99, mx = rel_err(f, exp)
    print("config1 FP32 ref-mode force error: global %.3e p99 %.3e max %.3e" % (glob, p99, mx))
    assert glob <= 1e-5
    assert mx <= 1e-3
    np.testing.assert_array_equal(pairs, oracle.detect(st))


@pytest.mark.parametrize("variant", [1, 2])
@pytest.mark.parametrize("ipl", [1, 2, 4])
def test_fp32_shapes_agree(oracle, variant, ipl):
    b = gen.plummer(5000, seed=3, radius=0.03)
    st = oracle.State.from_bodies(b)
    exp = oracle.force_soft(st, 1.0, 0.01, cutoff=True, nthreads=NT)
    with NBodyGPU(G=1.0, eps=0.01, variant=variant, ipl=ipl, tile_j=128, block=128,
                  jsplit=3) as g:
        g.load(b)
        f = g.forces()
        pairs = g.pairs()
        assert g.config() == {"ipl": ipl, "tile_j": 128, "jsplit": 3, "block": 128,
                              "variant": variant}
    glob, _, p99, _ = rel_err(f, exp)
    assert glob <= 1e-5 and p99 <= 1e-4
    np.testing.assert_array_equal(pairs, oracle.detect(st))
    assert len(pairs) > 100


def test_dense_collisions_overflow_flag_slots(oracle):
    """Radii so large that most targets flag more than 4 tiles: the fix-up recomputes whole
    slices in FP64; pairs and forces still match."""
    b = c1(n=3000, radius_k=0.25)
    st = oracle.State.from_bodies(b)
    exp = oracle.force_ref(st, G_REF, nthreads=NT)
    with NBodyGPU(tile_j=64, block=64, jsplit=2, pair_capacity=1 << 20) as g:
        g.load(b)
        f = g.forces()
        pairs = g.pairs()
    ref_pairs = oracle.detect(st)
    assert len(ref_pairs) > 20000
    np.testing.assert_array_equal(pairs, ref_pairs)
    glob, _, _, mx = rel_err(f, exp)
    assert glob <= 1e-5


def test_boundary_pairs_exact(oracle):
    """d == R1 + R2 exactly is a hit, a hair further is not, coincident bodies are a hit."""
    b = c1(n=2000, radius_k=0.01)
    b.x[:6] = [0.0, 3.0, 100.0, 103.0, 200.0, 200.0]
    b.y[:6] = [0.0, 4.0, 0.0, 4.0, 0.0, 0.0]
    b.z[:6] = [2000.0] * 6
    # 2.5 + nextafter(2.5, 0) rounds back to 5.0 (a tie, to even); two ulps below does not
    r3 = np.nextafter(np.nextafter(2.5, 0), 0)
    b.r[:6] = [2.5, 2.5, 2.5, r3, 1.0, 1.0]
    assert 2.5 + r3 < 5.0
    st = oracle.State.from_bodies(b)
    for prec in ("fp32", "fp64"):
        with NBodyGPU(precision=prec) as g:
            g.load(b)
            g.forces()
            pairs = {tuple(p) for p in g.pairs()}
        assert (0, 1) in pairs and (2, 3) not in pairs and (4, 5) in pairs
        assert pairs == {tuple(p) for p in oracle.detect(st)}


def test_fp32_steps_match_oracle_step_by_step(oracle):
    """Config 1 with collisions, 30 cycles: each GPU cycle against one oracle cycle from the same
    state: identical pairs, positions within FP32-force error."""
    b = c1()
    with NBodyGPU() as g:
        g.load(b)
        npairs = 0
        for _ in range(30):
            cur = g.store()
            st = oracle.State.from_bodies(cur)
            g.step(1e-3, 1.0)
            o = oracle.step(st, G_REF, 0.0, True, True, 1e-3, 1.0, nthreads=NT)
            np.testing.assert_array_equal(g.pairs(), o["pairs"])
            npairs += len(o["pairs"])
            nxt = g.store()
            np.testing.assert_allclose(nxt.pos(), st.pos(), rtol=1e-13, atol=1e-9)
            np.testing.assert_allclose(nxt.vel(), st.vel(), rtol=1e-6, atol=1e-9)
    assert npairs > 50


def test_config2_elastic_run(oracle):
    """Config 2 with elastic collisions (radius 1e-4), 50 steps: pair sets of the last cycle
    match the oracle's detection on the GPU state."""
    b = gen.plummer(65536, seed=21, radius=1e-4)
    with NBodyGPU(G=1.0, eps=0.01) as g:
        g.load(b)
        s = g.step(1e-3, 1.0, nsteps=49)
        cur = g.store()
        g.step(1e-3, 1.0)
        st = oracle.State.from_bodies(cur)
        np.testing.assert_array_equal(g.pairs(), oracle.detect(st))
    assert s["n_active"] == 65536


def test_diagnostics_match_oracle(oracle):
    b = gen.plummer(4000, seed=4)
    st = oracle.State.from_bodies(b)
    e = oracle.energy(st, 1.0, 0.01, nthreads=NT)
    with NBodyGPU(G=1.0, eps=0.01) as g:
        g.load(b)
        d = g.diagnostics()
    assert d["kinetic"] == pytest.approx(e["kinetic"], rel=1e-12)
    assert d["potential"] == pytest.approx(e["potential"], rel=1e-12)
    np.testing.assert_allclose(d["momentum"], e["momentum"], atol=1e-14)
    assert d["n_active"] == 4000 and d["mass"] == pytest.approx(1.0, rel=1e-12)


# ---------------------------------------------------------------------------------------------
# Edge cases the reference's tests and quirks pin
# ---------------------------------------------------------------------------------------------
def test_wp_compute_geometry(golden):
    """TestWpCompute (workpool_test.go:41-56) through the C-ABI."""
    from decimal import Decimal as D
    w = golden["wp_compute"]
    p = np.array(w["pos"], float)
    from nbodygo_amd import Bodies
    b = Bodies(p[:, 0], p[:, 1], p[:, 2], [0, 0], [0, 0], [0, 0], w["m"], w["r"])
    exp = float(D(w["f0"][0]))
    for prec, tol in (("fp64", 4e-16), ("fp32", 1e-6)):
        with NBodyGPU(precision=prec) as g:
            g.load(b)
            f = g.forces()
            assert f[0, 0] == pytest.approx(exp, rel=tol)
            assert f[1, 0] == pytest.approx(-exp, rel=tol)
            g.step(1.0, 1.0)
            out = g.store()
            assert out.vx[0] == pytest.approx(exp, rel=tol) and out.vx[1] != 0


def test_empty_and_single_body():
    from nbodygo_amd import Bodies
    with NBodyGPU() as g:
        z = np.zeros(0)
        g.load(Bodies(z, z, z, z, z, z, z, z))
        s = g.step(1e-3, 1.0, nsteps=3)
        assert s["interactions"] == 0 and len(g.pairs()) == 0
        one = np.ones(1)
        g.load(Bodies(one, one, one, one, one, one, one, one))
        g.step(0.5, 1.0)
        out = g.store()
        assert out.x[0] == 1.5 and out.vx[0] == 1.0


def test_ragged_sizes(oracle):
    for n in (1023, 1025, 4097):
        b = c1(n=n, seed=n)
        st = oracle.State.from_bodies(b)
        with NBodyGPU(precision="fp64") as g:
            g.load(b)
            np.testing.assert_array_equal(g.forces(), oracle.force_ref(st, G_REF, nthreads=NT))
        with NBodyGPU() as g:
            g.load(b)
            glob = rel_err(g.forces(), oracle.force_ref(st, G_REF, nthreads=NT))[0]
            assert glob <= 1e-5
            np.testing.assert_array_equal(g.pairs(), oracle.detect(st))


def test_inactive_bodies_are_ignored(oracle):
    b = c1(n=1500)
    b.flags[::7] = 0
    b.flags[3::11] |= L.NB_F_FRAGMENTING
    st = oracle.State.from_bodies(b)
    with NBodyGPU(precision="fp64") as g:
        g.load(b)
        np.testing.assert_array_equal(g.forces(), oracle.force_ref(st, G_REF, nthreads=NT))
        np.testing.assert_array_equal(g.pairs(), oracle.detect(st))
    with NBodyGPU() as g:
        g.load(b)
        f = g.forces()
        act = b.active()
        assert np.all(f[~act] == 0)
        assert rel_err(f[act], oracle.force_ref(st, G_REF, nthreads=NT)[act])[0] <= 1e-5


def test_collide_geometry_removes_nan_bodies():
    """TestCollide (body_collection_test.go:319-344) through a full cycle: coincident bodies are
    resolved to NaN velocities and removed by the NaN guard of Update."""
    b = c1(n=2000, radius_k=0.001)
    b.x[[5, 9]] = 500.0
    b.y[[5, 9]] = 500.0
    b.z[[5, 9]] = 500.0
    for prec in ("fp32", "fp64"):
        with NBodyGPU(precision=prec) as g:
            g.load(b)
            s = g.step(1e-3, 1.0)
            out = g.store()
        assert s["removed_total"] == 2 and s["n_active"] == 1998
        assert (out.flags[[5, 9]] & L.NB_F_EXISTS).tolist() == [0, 0]


def test_zero_mass_body_is_removed():
    b = c1(n=1200)
    b.m[17] = 0.0
    for prec in ("fp32", "fp64"):
        with NBodyGPU(precision=prec, collisions=False) as g:
            g.load(b)
            s = g.step(1e-3, 1.0)
        assert s["removed_total"] == 1


def test_pair_overflow_is_reported():
    b = c1(n=2000, radius_k=0.5)
    with NBodyGPU(pair_capacity=16) as g:
        g.load(b)
        with pytest.raises(NBodyError) as e:
            g.step(1e-3, 1.0)
        assert e.value.code == L.NB_E_PAIR_OVERFLOW
        s = g.step(1e-3, 1.0, allow_overflow=True)
        assert s["pair_overflow"] == 1 and s["pairs_last"] > 16


def test_autotune_keeps_results(oracle):
    b = gen.plummer(16384, seed=6, radius=0.002)
    st = oracle.State.from_bodies(b)
    with NBodyGPU(G=1.0, eps=0.01) as g:
        g.load(b)
        f0, p0 = g.forces(), g.pairs()
        cfg = g.autotune(reps=2)
        f1, p1 = g.forces(), g.pairs()
        assert g.config() == cfg
    np.testing.assert_array_equal(p0, p1)
    np.testing.assert_array_equal(p1, oracle.detect(st))
    assert np.linalg.norm(f1 - f0) / np.linalg.norm(f0) < 1e-6


def test_render_and_store_agree():
    b = gen.plummer(3000, seed=7)
    with NBodyGPU(G=1.0, eps=0.01, collisions=False, overlap_cutoff=False) as g:
        g.load(b)
        g.step(1e-3, 1.0, nsteps=5)
        xyz, ex = g.render()
        out = g.store()
    np.testing.assert_array_equal(xyz, out.pos().astype(np.float32))
    assert ex.all()


def test_full_size_force_is_momentum_conserving():
    """Config 4 size (N = 262,144): sum_i m_i a_i = 0 (Newton's third law) to FP32 accuracy."""
    b = gen.plummer(262144, seed=8)
    with NBodyGPU(G=1.0, eps=0.01, collisions=False, overlap_cutoff=False) as g:
        g.load(b)
        f = g.forces()
    assert np.abs(f.sum(axis=0)).max() < 1e-5 * np.abs(f).sum(axis=0).max()
