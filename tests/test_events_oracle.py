"""Pins of the events oracle (oracle/events.py): subsume and fragmentation as the reference has
them (body.go:171-186, 228-244; fragcalc.go:24-117), with hand-worked outcomes."""
import numpy as np

from nbodygo_amd.sim import Bodies
from oracle import events as EV

G = 1.0


def bodies(pos, m, r, beh, vel=None, frag=None, fstep=None):
    pos = np.asarray(pos, float)
    n = len(pos)
    vel = np.zeros((n, 3)) if vel is None else np.asarray(vel, float)
    return Bodies(pos[:, 0], pos[:, 1], pos[:, 2], vel[:, 0], vel[:, 1], vel[:, 2],
                  np.asarray(m, float), np.asarray(r, float), beh=np.asarray(beh, np.int32),
                  frag=frag, fstep=fstep)


def test_generator_is_uniform_and_draws_in_the_ball():
    u = np.array([EV.uniform(7, k // 3, k % 3) for k in range(60000)])
    assert 0 <= u.min() and u.max() < 1
    assert abs(u.mean() - 0.5) < 0.005 and abs(u.var() - 1 / 12) < 0.002
    v = np.array([EV.vector_even(11, k) for k in range(20000)])
    rad = np.linalg.norm(v, axis=1)
    assert rad.max() <= 1.0
    # uniform in the ball: P(|v| < s) = s^3
    for s in (0.25, 0.5, 0.75):
        assert abs(np.mean(rad < s) - s ** 3) < 0.01
    assert np.abs(v.mean(0)).max() < 0.02


def test_subsume_chain_in_pair_order():
    # A (Subsume, R 1) reaches B's centre (0.8 <= 1): A absorbs B.  A does not reach C
    # (1.1 > 1).  B (Subsume, R 0.5) reaches C (0.3 <= 0.5): in pair order B absorbs C after B
    # was absorbed, so C's mass goes to the removed B (no existence check in ResolveSubsume).
    b = bodies([(0, 0, 0), (0.8, 0, 0), (1.1, 0, 0), (50, 0, 0)], [5.0, 1.0, 2.0, 1.0],
               [1.0, 0.5, 0.2, 0.1], [EV.SUBSUME, EV.SUBSUME, EV.ELASTIC, EV.ELASTIC])
    sim = EV.EventSim(b, G, eps=0.01, cutoff=False)
    pairs = sim.step(1e-3, 1.0)
    assert pairs.tolist() == [[0, 1], [1, 2]]
    assert sim.a["m"].tolist() == [6.0, 2.0, 0.0, 1.0]
    assert (sim.flags & EV.EXISTS).tolist() == [1, 0, 0, 1]
    assert sim.stats["subsumed"] == 2
    assert sim.cycle(1.0) == 2


def test_equal_radii_and_none_behaviour():
    # equal radii: no subsume either way; None + Elastic: no event; Subsume + None: the larger
    # None body absorbs the Subsume one
    b = bodies([(0, 0, 0), (0.1, 0, 0), (10, 0, 0), (10.1, 0, 0), (20, 0, 0), (20.05, 0, 0)],
               [1, 1, 1, 1, 3, 1], [0.5, 0.5, 0.5, 0.2, 0.5, 0.1],
               [EV.SUBSUME, EV.SUBSUME, EV.NONE, EV.ELASTIC, EV.NONE, EV.SUBSUME])
    sim = EV.EventSim(b, G, eps=0.01, cutoff=False)
    pairs = sim.step(1e-3, 1.0)
    assert pairs.tolist() == [[4, 5]]
    assert sim.a["m"][4] == 4.0 and not (sim.flags[5] & EV.EXISTS)


def head_on(frag, fstep):
    # a Fragment body at rest hit by an Elastic one whose velocity components nearly cancel:
    # the factor dv / |vx + vy + vz| is large
    return bodies([(0, 0, 0), (0.15, 0, 0)], [1.0, 1.0], [0.1, 0.1],
                  [EV.FRAGMENT, EV.ELASTIC],
                  vel=[(0, 0, 0), (-1.0, 0.999, 0.0)], frag=[frag, 0.0], fstep=[fstep, 0.0])


def run_frag(frag, fstep, cycles, cycle_each=True):
    """Cycles of the head-on scene; with cycle_each=False the fragments stay queued until a
    last Cycle (so they take no part in the steps)."""
    sim = EV.EventSim(head_on(frag, fstep), G, eps=0.01, cutoff=False)
    per_cycle = []
    for _ in range(cycles):
        before = sim.stats["emitted"]
        sim.step(1e-3, 1.0)
        per_cycle.append(sim.stats["emitted"] - before)
        if cycle_each:
            sim.cycle(1.0)
    if not cycle_each:
        sim.cycle(1.0)
    return sim, per_cycle


def test_frag_factor_above_ten_caps_the_count():
    # FragFactor > 10: fragDelta = 10 whatever the impact, so FragStep 0.5 gives 5 fragments
    sim, per = run_frag(11.0, 0.5, 4)
    assert sim.stats["fragmentations"] == 1
    assert per == [0, 5, 0, 0]               # created the cycle after it starts
    assert sim.n == 1 + 5                     # the parent is gone, the fragments stay
    frags = sim.a["m"][1:]
    assert np.all(frags == 1.0 / 5.0) and np.all(sim.a["r"][1:] == 1.0)
    assert np.all(sim.beh[1:] == EV.ELASTIC)


def test_fragment_counts_truncate_and_the_threshold():
    sim, per = run_frag(11.0, 0.15, 3)       # 1.5 fragments: int(1.5) = 1
    assert sum(per) == 1 and sim.a["m"][-1] == 1.0 / 1.5
    sim, per = run_frag(11.0, 0.1, 3)        # 1.0 <= 1: no fragmentation, behaviour kept
    assert sum(per) == 0 and sim.stats["fragmentations"] == 0 and sim.n == 2
    assert sim.beh[0] == EV.FRAGMENT


def test_two_thousand_fragments_over_twenty_cycles():
    sim = EV.EventSim(head_on(11.0, 300.0), G, eps=0.01, cutoff=False)
    sim.step(1e-3, 1.0)                      # min(10 * 300, 2000) = 2000 to make
    centre = sim.fs_pos[0].copy()
    per = []
    for _ in range(21):
        before = len(sim.pending)
        sim.step(1e-3, 1.0)
        per.append(len(sim.pending) - before)
    assert per == [101] * 19 + [81, 0]
    assert not (sim.flags[0] & EV.EXISTS)
    # every fragment within 0.9 R of where the parent was when it started, at its velocity
    p = np.array([q[:3] for q in sim.pending])
    assert np.linalg.norm(p - centre, axis=1).max() <= 0.1 * 0.9 * (1 + 1e-12)
    assert len({q[3:6] for q in sim.pending[:101]}) == 1
    assert sim.cycle(1.0) == 1 + 2000


def test_fragmenting_body_keeps_its_force():
    # a fragmenting body is kicked each cycle by the force it had when it started
    sim = EV.EventSim(head_on(11.0, 1000.0), G, eps=0.01, cutoff=False)
    sim.step(1e-3, 1.0)
    assert sim.flags[0] & EV.FRAGMENTING
    f0 = sim.fs_f[0].copy()
    v0 = np.array([sim.a["vx"][0], sim.a["vy"][0], sim.a["vz"][0]])
    for k in range(1, 4):
        sim.step(1e-3, 1.0)
        v = np.array([sim.a["vx"][0], sim.a["vy"][0], sim.a["vz"][0]])
        np.testing.assert_allclose(v, v0 + k * 1e-3 * f0 / 1.0, rtol=1e-14, atol=1e-18)
