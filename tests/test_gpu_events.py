"""GPU tests of the subsume and fragmentation events (nb_params.events = 1) against the events
oracle (oracle/events.py).

Bars: event pair lists identical to the oracle's on every pair path (FP64, FP32 + FP64 fix-up,
Barnes-Hut, surrogate near field); flags, behaviours, body counts and event counts exact through
cycles of subsumes and fragmentation; masses to 1e-12 (a fragment's mass divides by
fragDelta * FragStep, and fragDelta comes from coll3d, whose transcendentals differ in the last
ulp between ocml and glibc), positions and velocities to 1e-10 relative.
"""
import numpy as np
import pytest

from nbodygo_amd import NBodyGPU, gen
from nbodygo_amd import _lib as L
from nbodygo_amd.sim import Bodies
from oracle import events as EV

pytestmark = pytest.mark.gpu

EPS = 0.01


def scene(n, seed, p_sub=0.3, p_frag=0.2, p_none=0.1, radius_k=0.25, fstep=3.0):
    """A uniform cluster whose radii make many overlaps, with every behaviour."""
    b = gen.uniform_sphere(n, seed=seed, radius_k=radius_k)
    rng = np.random.default_rng(seed)
    u = rng.random(n)
    b.beh[:] = L.NB_CB_ELASTIC
    b.beh[u < p_sub] = L.NB_CB_SUBSUME
    b.beh[(u >= p_sub) & (u < p_sub + p_frag)] = L.NB_CB_FRAGMENT
    b.beh[(u >= p_sub + p_frag) & (u < p_sub + p_frag + p_none)] = L.NB_CB_NONE
    b.frag[:] = rng.uniform(0.0, 2.0, n)
    b.fstep[:] = fstep
    b.vx += rng.normal(0, 0.3, n)
    b.vy += rng.normal(0, 0.3, n)
    return b


def gpu_ctx(**kw):
    kw.setdefault("precision", "fp64")
    return NBodyGPU(G=1.0, eps=EPS, overlap_cutoff=False, collisions=True, events=True,
                    pair_capacity=1 << 20, **kw)


def compare(g_state, sim, rtol=1e-10):
    assert g_state.n == sim.n
    np.testing.assert_array_equal(g_state.flags, sim.flags)
    np.testing.assert_array_equal(g_state.beh, sim.beh)
    # a fragment's mass is m / (fragDelta * FragStep), fragDelta from coll3d's outcome
    np.testing.assert_allclose(g_state.m, sim.a["m"], rtol=1e-12, atol=0)
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        np.testing.assert_allclose(getattr(g_state, k), sim.a[k], rtol=rtol, atol=1e-12,
                                   err_msg=k)


def test_cycles_with_subsumes_and_fragmentation_match_oracle():
    b = scene(1500, seed=1)
    sim = EV.EventSim(b, 1.0, eps=EPS, cutoff=False)
    with gpu_ctx() as g:
        g.load(b)
        totals = dict(sub=0, frag=0)
        for k in range(6):
            exp_pairs = sim.step(1e-3, 1.0)
            g.step(1e-3, 1.0, 1)
            np.testing.assert_array_equal(g.pairs(), exp_pairs, err_msg="cycle %d" % k)
            ev = g.event_stats()
            assert ev["subsumed_total"] == sim.stats["subsumed"]
            assert ev["fragmentations"] == sim.stats["fragmentations"]
            assert ev["fragments_pending"] == len(sim.pending)
            assert g.cycle(1.0) == sim.cycle(1.0)
            compare(g.store(), sim)
        assert sim.stats["subsumed"] > 20 and sim.stats["fragmentations"] > 5
        assert sim.stats["emitted"] > 20


def test_fragments_are_exact_and_the_parent_keeps_its_force():
    # the head-on scene of the oracle's pins: the fragments' positions, velocities and masses
    b = Bodies(np.array([0.0, 0.15, 5.0]), np.zeros(3), np.zeros(3),
               np.array([0.0, -1.0, 0.0]), np.array([0.0, 0.999, 0.0]), np.zeros(3),
               np.array([1.0, 1.0, 0.5]), np.array([0.1, 0.1, 0.1]),
               beh=np.array([L.NB_CB_FRAGMENT, L.NB_CB_ELASTIC, L.NB_CB_ELASTIC], np.int32),
               frag=np.array([11.0, 0.0, 0.0]), fstep=np.array([30.0, 0.0, 0.0]))
    sim = EV.EventSim(b, 1.0, eps=EPS, cutoff=False)
    with gpu_ctx() as g:
        g.load(b)
        for k in range(4):  # 300 fragments: 101, 101, 98 over three cycles after the hit
            sim.step(1e-3, 1.0)
            g.step(1e-3, 1.0, 1)
            assert g.cycle(1.0) == sim.cycle(1.0)
        s = g.store()
        ev = g.event_stats()
    assert sim.stats["emitted"] == ev["fragments_emitted"] == 300
    assert s.n == 2 + 300 and ev["fragmenting"] == 0
    compare(s, sim, rtol=1e-12)


@pytest.mark.parametrize("path", ["fp32", "barnes_hut", "surrogate"])
def test_event_pairs_on_every_path(path):
    b = scene(6000, seed=2, radius_k=0.1)
    if path == "surrogate":
        b = gen.two_plummer(6000, seed=3, radius_k=0.5)
        b.beh[::3] = L.NB_CB_SUBSUME
        b.beh[1::7] = L.NB_CB_NONE
    exp = EV.EventSim(b, 1.0, eps=EPS, cutoff=False).detect()
    assert len(exp) > 30
    kw = dict(precision="fp64") if path == "barnes_hut" else dict(precision="fp32")
    with gpu_ctx(**kw) as g:
        g.load(b)
        if path == "barnes_hut":
            g.bh_enable()
        elif path == "surrogate":
            g.sur_enable(check_every=0, use_mlp=False)
        g.forces()
        np.testing.assert_array_equal(g.pairs(), exp)


def test_events_off_keeps_the_elastic_only_path():
    b = scene(1500, seed=4)
    with NBodyGPU(G=1.0, eps=EPS, overlap_cutoff=False, collisions=True,
                  precision="fp64") as g:
        g.load(b)
        g.step(1e-3, 1.0, 1)
        s = g.store()
        assert g.pairs().shape[0] > 0
    assert np.all(s.m == b.m)  # no subsume
    assert not np.any(s.flags & L.NB_F_FRAGMENTING)
