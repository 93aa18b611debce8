"""The reference's own collection tests (cmd/body/body_collection_test.go) re-expressed through
the C-ABI on the device.  Their bodies come from createTestBody (:26-56): every float field of
body `id` equals id, Elastic, Exists, restitution 1.  The device keeps no ids, so a body is
recognised by its position x = id."""
import numpy as np
import pytest

from nbodygo_amd import NBodyGPU
from nbodygo_amd import _lib as L
from nbodygo_amd.sim import Bodies

pytestmark = pytest.mark.gpu


def reference_bodies(cnt):
    v = np.arange(cnt, dtype=np.float64)
    return Bodies(*(v.copy() for _ in range(8)), frag=v.copy(), fstep=v.copy())


def test_init_size():
    """TestInitSize (:59-65): a collection of 1,000,000 bodies has that count."""
    with NBodyGPU() as g:
        g.load(reference_bodies(1_000_000))
        assert g.n == 1_000_000
        out = g.store()
    assert out.n == 1_000_000 and np.array_equal(out.x, np.arange(1_000_000, dtype=np.float64))


def test_remove():
    """TestRemove (:71-88): body 10 set to not exist, then Cycle removes it."""
    with NBodyGPU() as g:
        g.load(reference_bodies(1000))
        g.set_body(10, flags=0)
        assert g.cycle(1.0) == 999
        out = g.store()
    assert 10.0 not in out.x
    assert np.array_equal(out.x, np.delete(np.arange(1000, dtype=np.float64), 10))


def test_adds():
    """TestAdds (:266-281): an add of body 600 to 500 bodies is picked up by Cycle."""
    with NBodyGPU() as g:
        g.load(reference_bodies(500))
        add = reference_bodies(601)
        one = Bodies(add.x[600:], add.y[600:], add.z[600:], add.vx[600:], add.vy[600:],
                     add.vz[600:], add.m[600:], add.r[600:], frag=add.frag[600:],
                     fstep=add.fstep[600:])
        g.add_bodies(one)
        assert g.cycle(1.0) == 501
        b = g.get_body(500)
    assert b["x"] == 600.0 and b["m"] == 600.0 and b["flags"] & L.NB_F_EXISTS


def test_subsume():
    """TestSubsume (:283-313): body 10 subsumes body 444; afterwards 10 exists and 444 does
    not.  The reference enqueues that event by hand; here it comes from Compute's own dispatch
    (body.go:178-183): 444 is put inside 10, 10 is a Subsume body with the larger radius, and one
    cycle with events on resolves it (ResolveSubsume, body.go:228-244)."""
    b = reference_bodies(500)
    b.x[:] = np.arange(500) * 10.0  # apart: no other pair overlaps
    b.y[:] = b.z[:] = 0.0
    b.vx[:] = b.vy[:] = b.vz[:] = 0.0
    b.r[:] = 1.0
    b.x[444] = b.x[10] + 2.0
    b.r[10] = 3.0
    b.beh[10] = L.NB_CB_SUBSUME
    with NBodyGPU(G=1e-9, events=True) as g:
        g.load(b)
        g.step(1e-3, 1.0, 1)
        b10, b444 = g.get_body(10), g.get_body(444)
        st = g.event_stats()
    assert b10["flags"] & L.NB_F_EXISTS and not (b444["flags"] & L.NB_F_EXISTS)
    assert b10["m"] == 10.0 + 444.0
    assert st["subsumed_total"] == 1
