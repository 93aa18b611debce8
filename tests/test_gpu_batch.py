"""Batched path (nb_batch_*): many independent simulations in one persistent launch.

Bars: per-simulation forces within FP32 error of the FP64 oracle (global <= 1e-5); simulations
never see each other; the hipGraph launch is bit-identical to eager launches of the same kernels,
replay after replay; the target-block width (IPL) does not change any body's arithmetic, and
every (waves, tile) tuning meets the same oracle bar."""
import numpy as np
import pytest

from nbodygo_amd import NBodyBatch, NBodyGPU, gen

pytestmark = pytest.mark.gpu
NT = 16


def sims_mixed(seed=1):
    sizes = [300, 57, 1500, 0, 2048, 5000, 777, 64, 4096, 1]
    return gen.batch_clusters(sizes, seed=seed, radius=1.0, vel_sigma=0.05)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("waves,buckets", [
    (4, [(2, 64), (4, 64), (4, 64)]), (1, [(2, 256), (4, 256), (4, 256)]),
    (2, [(2, 128), (4, 128), (4, 128)]),
    # one shape for every bucket: the kernel specialised to it
    (4, [(1, 256)] * 3), (2, [(2, 64)] * 3), (1, [(4, 128)] * 3)])
def test_batch_forces_match_oracle_per_sim(oracle, waves, buckets):
    sims = sims_mixed()
    with NBodyBatch(G=1.0, eps=0.01) as bt:
        bt.load(sims)
        bt.set_tuning(waves, buckets)
        f = bt.forces()
    off = bt.offsets
    for k, s in enumerate(sims):
        if s.n < 2:
            assert np.all(f[off[k]:off[k + 1]] == 0)
            continue
        st = oracle.State.from_bodies(s)
        exp = oracle.force_soft(st, 1.0, 0.01, nthreads=NT)
        assert rel(f[off[k]:off[k + 1]], exp) <= 1e-5, k


def test_simulations_do_not_interact(oracle):
    a = gen.uniform_sphere(700, seed=3, radius=1.0, mass_mu=0.0, vel_sigma=0.1)
    b = gen.uniform_sphere(900, seed=4, radius=1.0, mass_mu=0.0, vel_sigma=0.1)
    with NBodyBatch(G=1.0, eps=0.01) as bt:
        bt.load([a, b])
        f = bt.forces()
    for s, part in ((a, f[:700]), (b, f[700:])):
        exp = oracle.force_soft(oracle.State.from_bodies(s), 1.0, 0.01, nthreads=NT)
        assert rel(part, exp) <= 1e-5


def test_graph_replays_match_eager_launches():
    sims = sims_mixed(seed=2)
    with NBodyBatch(G=1.0, eps=0.01) as g1, NBodyBatch(G=1.0, eps=0.01) as g2:
        g1.load(sims)
        g2.load(sims)
        g2.set_graph(False)
        for rep in range(4):
            g1.step(1e-3, 1.0, nsteps=5)      # graph captured once, then replayed
            g2.step(1e-3, 1.0, nsteps=5)      # eager launches of the same kernels
            f_eager = g2.forces()             # eager work between replays
            s1, s2 = g1.store(), g2.store()
            for k in ("x", "y", "z", "vx", "vy", "vz"):
                np.testing.assert_array_equal(getattr(s1, k), getattr(s2, k))
            np.testing.assert_array_equal(g1.forces(), f_eager)


def test_shapes_do_not_change_results():
    """IPL only regroups targets: forces are bit-identical across it for a given (waves, tile).
    Waves and tile change where the Kahan tiles and the source split fall, so forces agree to
    FP32 rounding (1e-6 relative), and every tuning gives the oracle's forces to the FP32 bar."""
    sims = sims_mixed(seed=5)
    out = []
    for ipl in (1, 2, 4):
        with NBodyBatch(G=1.0, eps=0.01) as bt:
            bt.load(sims)
            bt.set_shapes([ipl, ipl, ipl])
            assert bt.shapes() == [ipl, ipl, ipl]
            out.append(bt.forces())
    np.testing.assert_array_equal(out[0], out[1])
    np.testing.assert_array_equal(out[0], out[2])
    # one IPL per bucket runs the kernel specialised to it; mixed IPLs the per-item dispatch
    with NBodyBatch(G=1.0, eps=0.01) as bt:
        bt.load(sims)
        bt.set_shapes([1, 2, 4])
        np.testing.assert_array_equal(bt.forces(), out[0])
    for waves in (1, 2, 4):
        for tile in (64, 128, 256):
            with NBodyBatch(G=1.0, eps=0.01) as bt:
                bt.load(sims)
                bt.set_tuning(waves, [(2, tile), (4, tile), (1, tile)])
                assert bt.tuning() == {"waves": waves, "buckets": [(2, tile), (4, tile), (1, tile)]}
                f = bt.forces()
            assert rel(f, out[0]) < 1e-6, (waves, tile)
    with NBodyBatch(G=1.0, eps=0.01) as bt:
        bt.load(sims)
        tuned = bt.autotune(reps=1)
        assert tuned["waves"] in (1, 2, 4)
        assert all(i in (1, 2, 4) and t in (64, 128, 256) for i, t in tuned["buckets"])
        assert rel(bt.forces(), out[0]) < 1e-6


def test_batch_step_tracks_single_context():
    # total mass ~0.5 in a unit ball: smooth trajectories over 20 steps, so the two paths' FP32
    # force differences (different tile partitions) stay at rounding level
    s = gen.uniform_sphere(3000, seed=6, radius=1.0, mass_mu=np.log(1e-4), vel_sigma=0.1)
    with NBodyBatch(G=1.0, eps=0.01) as bt, NBodyGPU(G=1.0, eps=0.01, collisions=False,
                                                     overlap_cutoff=False) as g:
        bt.load([s])
        g.load(s)
        assert rel(bt.forces(), g.forces()) < 1e-6
        bt.step(1e-3, 1.0, nsteps=20)
        g.step(1e-3, 1.0, nsteps=20)
        a, b = bt.store(), g.store()
    dv = np.abs(s.vel() - b.vel()).max()  # velocity change over the run
    assert np.abs(a.vel() - b.vel()).max() <= 1e-5 * dv
    assert np.abs(a.pos() - b.pos()).max() <= 1e-5 * dv * 20e-3


def test_config3_shape_momentum_per_sim():
    """Config 3 sizes (512 simulations, N uniform in [256, 8192]): each simulation's total force
    vanishes (Newton's third law) to FP32 accuracy."""
    sizes = gen.batch_sizes(512, seed=7)
    sims = gen.batch_clusters(sizes, seed=7)
    with NBodyBatch(G=1.0, eps=0.01) as bt:
        bt.load(sims)
        st = bt.step(1e-3, 1.0, nsteps=2)
        f = bt.forces()
    assert st["interactions"] == 2 * sum(int(n) * (int(n) - 1) for n in sizes)
    off = bt.offsets
    for k in range(len(sizes)):
        part = f[off[k]:off[k + 1]]
        assert np.abs(part.sum(axis=0)).max() <= 1e-5 * np.abs(part).sum(axis=0).max() + 1e-300
