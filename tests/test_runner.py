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
import os
import subprocess

import numpy as np
import pytest

from nbodygo_amd import _lib as L
from nbodygo_amd import csvio, gen
from nbodygo_amd import runner as R

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


def test_delete_rule_mirrors_the_reference_walk():
    ex = np.ones(10, bool)
    pin = np.zeros(10, bool)
    # step = 10 // 3 = 3: stride points 0, 3, 6, 9; the fourth removal is the closure quirk
    assert R.remove_indices(ex, pin, 3) == [0, 3, 6, 9]
    pin[3] = True  # a pinned body passes the removal on to the next live body
    assert R.remove_indices(ex, pin, 3) == [0, 4, 6, 9]
    ex[0] = False
    assert R.remove_indices(ex, pin, 3) == [1, 4, 6, 9]
    assert R.remove_indices(ex, pin, 50) == [1, 2, 4, 5, 6, 7, 8, 9]  # more than exist: step 1
    assert R.remove_indices(ex, pin, -1) == [1, 2, 3, 4, 5, 6, 7, 8, 9]
    assert R.remove_indices(ex, pin, 0) == []


def test_collection_cycle_compacts_and_appends():
    b = gen.uniform_sphere(6, seed=1)
    bc = R.BodyCollection(b)
    b.flags[[1, 4]] = 0
    bc.b.flags[[1, 4]] = 0
    ids = bc.ids.copy()
    add = gen.uniform_sphere(2, seed=2)
    bc.enqueue_add(add)
    assert bc.cycle(0.5)
    assert bc.count() == 6 and list(bc.ids[:4]) == list(ids[[0, 2, 3, 5]])
    assert np.all(bc.b.rest[4:] == 0.5)
    assert not bc.cycle(0.5)


# ---------------------------------------------------------------------------------------------
# GPU: both drivers against the oracle
# ---------------------------------------------------------------------------------------------
def oracle_run(oracle, b, steps, dt, removals=(), collisions=True):
    """Oracle state after `steps` cycles; removals = [(cycle, count)] applied before that cycle
    with the reference's delete rule on the (uncompacted) array."""
    st = oracle.State.from_bodies(b)
    for k in range(steps):
        for at, cnt in removals:
            if at == k:
                live = (st.flags & L.NB_F_EXISTS).astype(bool)
                # the drivers compact after each cycle, so the walk sees live bodies only
                order = np.nonzero(live)[0]
                idx = R.remove_indices(np.ones(order.size, bool),
                                       (st.flags[order] & L.NB_F_PINNED).astype(bool), cnt)
                for i in order[idx]:
                    st.flags[i] &= ~np.uint8(L.NB_F_EXISTS)
                    st.a["m"][i] = 0.0
        oracle.step(st, G_REF, 0.0, True, collisions, dt, 1.0, 1, nthreads=16)
    live = (st.flags & L.NB_F_EXISTS).astype(bool)
    return np.stack([st.x, st.y, st.z, st.vx, st.vy, st.vz], 1)[live]


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
    exp = oracle_run(oracle, b, 40, 1e-3, removals, collisions)
    assert_tracks(np.stack([got.x, got.y, got.z, got.vx, got.vy, got.vz], 1), exp, collisions)


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
    exp = oracle_run(oracle, b, 30, 1e-3, ((5, 9),), collisions)
    assert_tracks(got, exp, collisions)
    assert run.stats["computations"] == 30 and run.stats["loads"] >= 2


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
