"""The host drivers above the C-ABI: the C++ nb_server (nbodygo_amd/runner/) and the Python
ctypes runner (nbodygo_amd/runner.py), both in the reference's phase order
(computation-runner.go:267-326).

CPU: the C++ CSV loader against the Python one (fromcsv.go semantics), the delete rule of
processDeletes (computation-runner.go:176-216) with its closure quirk, compaction.
GPU: both drivers pinned to the oracle on the FP64 path (config 1, reference mode), including
removals and the compaction that follows them: bit for bit without collisions; with elastic
collisions to 1e-12 relative (coll3d's acos/atan2/sin/cos come from ocml on the device and glibc
in the oracle, so they may differ in the last ulp).
"""
import itertools
import os
import subprocess

import numpy as np
import pytest

from nbodygo_amd import _lib as L
from nbodygo_amd import csvio, gen
from nbodygo_amd import runner as R
from oracle import lifecycle as LC

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SERVER = os.path.join(ROOT, "nbodygo_amd", "nb_server")
G_REF = 6.673e-11

TRICKY = '''# a comment line
1,2,3,4,5,6,7,8

"1.5",2,3,4,5,6,7,8,true,subsume,Blue,0.5,2
1,2,3,4,5,6,7
x,1,2,3,4,5,6,7
1e400,1,1,1,1,1,1,1
0x1p-3,1,1,1,1,1,1,1,F
1,1,1,1,1,1,1,1,,elastic
  2 ,3,4,5,6,7,8,9,1,NONE,weird
-inf,1,1,1,1,1,1,1,0,fragment,pink,1e-3,0x1.8p1
"quoted ""name""",1,1,1,1,1,1,1
+.5e-3,-0.,1E+2,4,5,6,7,8,TRUE,Elastic,RED\r
7,7,7,7,7,7,7,7
8,8,8,8,8,8,8,8
'''


def parse_cpp(path, body_count):
    out = subprocess.run([SERVER, "-f", str(path), "-b", str(body_count), "--parse-only"],
                         capture_output=True, text=True)
    return out.returncode, [list(map(float, l.split())) for l in out.stdout.splitlines()], \
        out.stderr


@pytest.mark.parametrize("body_count", [1000, 5, 1, 0])
def test_cpp_csv_loader_matches_python(tmp_path, body_count):
    p = tmp_path / "b.csv"
    p.write_text(TRICKY)
    rc, rows, err = parse_cpp(p, body_count)
    assert rc == 0, err
    b, colors, suns = csvio.read_csv(str(p), body_count=body_count)
    assert len(rows) == b.n
    for i, r in enumerate(rows):
        exp = [b.x[i], b.y[i], b.z[i], b.vx[i], b.vy[i], b.vz[i], b.m[i], b.r[i], float(suns[i]),
               float(b.beh[i]), float(colors[i]), b.frag[i]]
        np.testing.assert_array_equal(np.array(r[:12]), np.array(exp))
    if body_count == 1000:
        assert b.n == 8  # not parsed: the 7-field, "x", 1e400, empty is_sun and quoted-name rows
        assert b.x[4] == -np.inf and b.frag[4] == 1e-3


def test_cpp_csv_loader_stops_before_a_later_malformed_record(tmp_path):
    p = tmp_path / "bad.csv"
    p.write_text('1,2,3,4,5,6,7,8\n1,2,"oops,4,5,6,7,8\n')
    rc, rows, err = parse_cpp(p, 1)
    assert rc == 0 and len(rows) == 1     # the reader never reaches the bad record
    rc, rows, err = parse_cpp(p, 1000)
    assert rc == 1 and "quoted-field" in err


# ---------------------------------------------------------------------------------------------
# GPU: both drivers against the oracle
# ---------------------------------------------------------------------------------------------
def oracle_run(oracle, b, steps, dt, removals=(), collisions=True, adds=(), mods=(), gets=()):
    """The runner's cycles on the oracle: before cycle k the removals [(k, count)] (the delete
    walk over the compacted array), the mods [(k, id, {field: value})] and gets [(k, id)]; then
    the step; then Cycle with the adds [(k, Bodies)] enqueued before cycle k.  Ids are 0..n-1
    for the loaded bodies and count on through each add in enqueue order.  Returns the final
    live state and the get answers (the body's row before that cycle's step)."""
    st = oracle.State.from_bodies(b)
    ids = np.arange(b.n)
    nid = b.n
    answers = {}
    for k in range(steps):
        for at, cnt in removals:
            if at == k:
                idx = LC.remove_indices((st.flags & L.NB_F_EXISTS).astype(bool),
                                        (st.flags & L.NB_F_PINNED).astype(bool), cnt)
                LC.apply_removals(st, idx)
        for at, id_, fields in mods:
            if at == k:
                for i in np.nonzero(ids == id_)[0]:
                    for f, v in fields.items():
                        getattr(st, f)[i] = v
        for at, id_ in gets:
            if at == k:
                i = np.nonzero(ids == id_)[0][0]
                answers[(k, id_)] = [getattr(st, f)[i] for f in
                                     ("x", "y", "z", "vx", "vy", "vz", "m", "r")]
        oracle.step(st, G_REF, 0.0, True, collisions, dt, 1.0, 1, nthreads=16)
        pend = [a for at, a in adds if at == k]
        keep = (st.flags & L.NB_F_EXISTS).astype(bool)
        if pend or not keep.all():
            new_ids = []
            for a in pend:
                new_ids.append(np.arange(nid, nid + a.n))
                nid += a.n
            ids = np.concatenate([ids[keep]] + [x[::-1] for x in new_ids[::-1]])
            out = LC.cycle(st, pend, 1.0)
            st = oracle.State(*(out[f] for f in ("x", "y", "z", "vx", "vy", "vz", "m", "r")),
                              beh=out["beh"], flags=out["flags"], rest=out["rest"],
                              frag=out["frag"])
    live = (st.flags & L.NB_F_EXISTS).astype(bool)
    return np.stack([st.x, st.y, st.z, st.vx, st.vy, st.vz], 1)[live], answers


def assert_tracks(got, exp, collisions):
    if collisions:
        np.testing.assert_allclose(got, exp, rtol=1e-12, atol=1e-12 * np.abs(exp).max())
    else:
        np.testing.assert_array_equal(got, exp)


@pytest.mark.gpu
@pytest.mark.parametrize("collisions,removals", [(False, ()), (False, ((10, 7),)),
                                                 (True, ()), (True, ((10, 7),))])
def test_cpp_driver_matches_oracle_fp64(oracle, tmp_path, collisions, removals):
    b = gen.uniform_sphere(1000, seed=11, radius_k=0.05)
    src = tmp_path / "c1.csv"
    csvio.write_csv(b, str(src))
    b, _, _ = csvio.read_csv(str(src))  # the bodies exactly as the CSV carries them
    out = tmp_path / "final.csv"
    cmd = [SERVER, "-f", str(src), "-b", "1000", "-m", "1e-3", "--iterations", "40", "--fp64",
           "--out", str(out)]
    for at, cnt in removals:
        cmd += ["--remove", "%d@%d" % (cnt, at)]
    if not collisions:
        cmd += ["--no-collisions"]
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    assert "computations: 40" in res.stdout
    got, _, _ = csvio.read_csv(str(out), body_count=10 ** 6)
    exp, _ = oracle_run(oracle, b, 40, 1e-3, removals, collisions)
    assert_tracks(np.stack([got.x, got.y, got.z, got.vx, got.vy, got.vz], 1), exp, collisions)


@pytest.mark.gpu
def test_cpp_driver_adds_mods_and_gets_match_oracle(oracle, tmp_path):
    """AddBody, ModBody and GetBody through nb_server: adds land in Cycle newest first, mods
    patch the device between cycles (EXISTS=false compacts the body away), gets read it."""
    b = gen.uniform_sphere(1000, seed=14, radius_k=0.05)
    src = tmp_path / "c1.csv"
    csvio.write_csv(b, str(src))
    b, _, _ = csvio.read_csv(str(src))
    a = gen.uniform_sphere(40, seed=15, radius_k=0.05)
    a.x += 0.5
    add = tmp_path / "add.csv"
    csvio.write_csv(a, str(add))
    a, _, _ = csvio.read_csv(str(add))
    out = tmp_path / "final.csv"
    cmd = [SERVER, "-f", str(src), "-b", "1000", "-m", "1e-3", "--iterations", "30", "--fp64",
           "--out", str(out), "--add", "%s@5" % add, "--remove", "13@9",
           "--mod", "3@12:X=0.25,VX=-0.5,MASS=2e3", "--mod", "1020@14:EXISTS=false",
           "--mod", "999999@14:X=1", "--get", "3@12", "--get", "1005@20", "--get", "7@25"]
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    assert "mod 3: 3" in res.stdout and "mod 1020: 3" in res.stdout
    assert "mod 999999: 0" in res.stdout
    got, _, _ = csvio.read_csv(str(out), body_count=10 ** 6)
    exp, ans = oracle_run(oracle, b, 30, 1e-3, ((9, 13),), True, adds=((5, a),),
                          mods=((12, 3, dict(x=0.25, vx=-0.5, m=2e3)),
                                (14, 1020, dict(flags=0))),
                          gets=((12, 3), (20, 1005), (25, 7)))
    assert_tracks(np.stack([got.x, got.y, got.z, got.vx, got.vy, got.vz], 1), exp, True)
    lines = {l.split(":")[0]: l.split(":")[1].split() for l in res.stdout.splitlines()
             if l.startswith("get ")}
    for (k, id_), row in ans.items():
        vals = lines["get %d" % id_]
        if vals == ["none"]:
            assert False, (k, id_)
        np.testing.assert_allclose(np.array(vals[:8], float), row, rtol=1e-12, atol=1e-15)


@pytest.mark.gpu
@pytest.mark.parametrize("collisions", [False, True])
def test_python_runner_matches_oracle_fp64(oracle, collisions):
    b = gen.uniform_sphere(800, seed=12, radius_k=0.05)
    rqh = R.ResultQueueHolder(4)
    bc = R.BodyCollection(b)
    run = R.ComputationRunner(1, 1e-3, rqh, bc, precision="fp64", collisions=collisions)
    try:
        for k in range(30):
            if k == 5:
                run.remove_bodies(9)
            run.run_one_computation()
            assert rqh.next() is not None
        run.sync_to_host()
        got = np.stack([bc.b.x, bc.b.y, bc.b.z, bc.b.vx, bc.b.vy, bc.b.vz], 1)
    finally:
        run.close()
    exp, _ = oracle_run(oracle, b, 30, 1e-3, ((5, 9),), collisions)
    assert_tracks(got, exp, collisions)
    assert run.stats["computations"] == 30 and run.stats["loads"] == 1
    assert run.stats["compactions"] >= 1


@pytest.mark.gpu
def test_python_runner_adds_mods_and_gets_match_oracle(oracle, monkeypatch):
    b = gen.uniform_sphere(800, seed=16, radius_k=0.05)
    a1 = gen.uniform_sphere(30, seed=17, radius_k=0.05)
    a2 = gen.uniform_sphere(20, seed=18, radius_k=0.05)
    a2.y -= 0.4
    rqh = R.ResultQueueHolder(4)
    bc = R.BodyCollection(b, ids=np.arange(800))
    monkeypatch.setattr(R, "_ids", itertools.count(800))  # ids go on as the oracle counts
    run = R.ComputationRunner(1, 1e-3, rqh, bc, precision="fp64", collisions=True)
    got_rows = {}
    try:
        for k in range(25):
            if k == 4:
                bc.enqueue_add(a1)
                bc.enqueue_add(a2)
            if k == 8:
                run.remove_bodies(11)
            if k == 10:
                assert run.mod_body(5, x=0.3, vz=0.2) == 1
                assert run.mod_body(845, flags=0) == 1
                assert run.mod_body(10 ** 7, x=1.0) == 0
            if k in (10, 20):
                for id_ in (5, 812):
                    g = run.get_body(id_)
                    got_rows[(k, id_)] = [getattr(g, f)[0] for f in
                                          ("x", "y", "z", "vx", "vy", "vz", "m", "r")]
            n = bc.count()
            run.run_one_computation()
            q = rqh.next()
            assert q is not None and len(q.queue) == n
        run.sync_to_host()
        got = np.stack([bc.b.x, bc.b.y, bc.b.z, bc.b.vx, bc.b.vy, bc.b.vz], 1)
    finally:
        run.close()
    exp, ans = oracle_run(oracle, b, 25, 1e-3, ((8, 11),), True, adds=((4, a1), (4, a2)),
                          mods=((10, 5, dict(x=0.3, vz=0.2)), (10, 845, dict(flags=0))),
                          gets=((10, 5), (10, 812), (20, 5), (20, 812)))
    assert_tracks(got, exp, True)
    for key, row in ans.items():
        np.testing.assert_allclose(got_rows[key], row, rtol=1e-12, atol=1e-15)
    assert run.stats["loads"] == 1


@pytest.mark.gpu
def test_python_runner_skips_cycles_when_the_consumer_lags():
    b = gen.uniform_sphere(300, seed=13)
    rqh = R.ResultQueueHolder(2)
    run = R.ComputationRunner(1, 1e-3, rqh, R.BodyCollection(b), precision="fp64")
    try:
        for _ in range(5):
            run.run_one_computation()
    finally:
        run.close()
    assert run.stats["computations"] == 2 and run.stats["skipped_no_queue"] == 3


@pytest.mark.gpu
def test_cpp_driver_with_events_matches_oracle(tmp_path):
    """nb_server --events: subsumes and fragmentation through the C++ runner's cycles (the
    device queue appended by nb_cycle, the host collection compacted alongside)."""
    from oracle import events as EV
    b = gen.uniform_sphere(600, seed=19, radius_k=0.25)
    rng = np.random.default_rng(19)
    u = rng.random(b.n)
    b.beh[u < 0.3] = L.NB_CB_SUBSUME
    b.beh[(u >= 0.3) & (u < 0.5)] = L.NB_CB_FRAGMENT
    b.frag[:] = rng.uniform(0.0, 2.0, b.n)
    b.fstep[:] = 3.0
    src = tmp_path / "ev.csv"
    csvio.write_csv(b, str(src))
    b, _, _ = csvio.read_csv(str(src))
    out = tmp_path / "final.csv"
    cmd = [SERVER, "-f", str(src), "-b", "600", "-m", "1e-3", "--iterations", "8", "--fp64",
           "--events", "--G", "1.0", "--eps", "0.01", "--no-cutoff", "--out", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    sim = EV.EventSim(b, 1.0, eps=0.01, cutoff=False)
    for _ in range(8):
        sim.step(1e-3, 1.0)
        sim.cycle(1.0)
    assert sim.stats["subsumed"] > 10 and sim.stats["emitted"] > 10
    got, _, _ = csvio.read_csv(str(out), body_count=10 ** 6)
    live = (sim.flags & EV.EXISTS) != 0
    exp = np.stack([sim.a[k][live] for k in ("x", "y", "z", "vx", "vy", "vz")], 1)
    assert got.n == exp.shape[0]
    np.testing.assert_allclose(np.stack([got.x, got.y, got.z, got.vx, got.vy, got.vz], 1), exp,
                               rtol=1e-10, atol=1e-9)
    np.testing.assert_allclose(got.m, sim.a["m"][live], rtol=1e-12)


@pytest.mark.gpu
def test_python_runner_with_events_matches_oracle():
    from oracle import events as EV
    b = gen.uniform_sphere(500, seed=20, radius_k=0.25)
    rng = np.random.default_rng(20)
    u = rng.random(b.n)
    b.beh[u < 0.3] = L.NB_CB_SUBSUME
    b.beh[(u >= 0.3) & (u < 0.5)] = L.NB_CB_FRAGMENT
    b.frag[:] = rng.uniform(0.0, 2.0, b.n)
    b.fstep[:] = 3.0
    rqh = R.ResultQueueHolder(4)
    bc = R.BodyCollection(b)
    run = R.ComputationRunner(1, 1e-3, rqh, bc, precision="fp64", collisions=True, events=True,
                              G=1.0, eps=0.01, overlap_cutoff=False)
    sim = EV.EventSim(b, 1.0, eps=0.01, cutoff=False)
    try:
        for _ in range(6):
            n = bc.count()
            run.run_one_computation()
            q = rqh.next()
            assert q is not None and len(q.queue) == n
            sim.step(1e-3, 1.0)
            sim.cycle(1.0)
            assert bc.count() == sim.n
        run.sync_to_host()
    finally:
        run.close()
    assert sim.stats["subsumed"] > 10 and sim.stats["emitted"] > 10
    assert len(set(bc.ids.tolist())) == bc.count()  # every body, fragments too, has its own id
    np.testing.assert_array_equal(bc.b.flags, sim.flags)
    np.testing.assert_allclose(bc.b.m, sim.a["m"], rtol=1e-12)
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        np.testing.assert_allclose(getattr(bc.b, k), sim.a[k], rtol=1e-10, atol=1e-9)
    np.testing.assert_array_equal(bc.b.r, sim.a["r"])  # fragments have radius 1
    assert (bc.b.r == 1.0).sum() > 10
