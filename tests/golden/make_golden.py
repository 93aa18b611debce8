"""Writes tests/golden/known_answers.json: known-answer vectors for the per-cycle path computed
from closed forms in 50-digit decimal arithmetic, independently of the oracle and of the GPU path.

Geometries re-expressed from the reference's own tests:
  wp_compute   TestWpCompute (cmd/runner/workpool_test.go:41-56): two bodies, m = 1, R = 1, at
               (1,1,1) and (22,22,22), G = 6.673e-11; force, then Update(1, 1).
  collide      TestCollide (cmd/body/body_collection_test.go:319-344): two coincident bodies;
               resolution marks both collided with NaN velocities (acos(0/0)).
Closed forms for the elastic collision (coll3d, collisioncalc.go:42-186, with restitution via
doElastic collisioncalc.go:26-35):
  head_on      equal masses, along the line of centres: velocities swap.
  one_d        unequal masses along the line of centres: v1' = (m1-m2)/(m1+m2) v1 + 2 m2/(m1+m2) v2.
  oblique      3-D impacts: the impulse acts along the line of centres at first contact,
               found by solving |x21 - u t| = r1 + r2 for the earlier root (u = v1 - v2).
Run: python tests/golden/make_golden.py
"""
import json
import os
import random
from decimal import Decimal as D, getcontext

getcontext().prec = 50
G_REF = D("6.673e-11")


def dstr(x):
    return str(+x)


def vsub(a, b):
    return [a[k] - b[k] for k in range(3)]


def vdot(a, b):
    return sum(a[k] * b[k] for k in range(3))


def wp_compute():
    p0 = [D(1), D(1), D(1)]
    p1 = [D(22), D(22), D(22)]
    d = vsub(p1, p0)
    dist = vdot(d, d).sqrt()
    F = G_REF / (dist * dist)
    f0 = [F * d[k] / dist for k in range(3)]
    # Update(1, 1): V += 1 * f / 1; X += 1 * V
    v0 = f0
    x0 = [p0[k] + v0[k] for k in range(3)]
    return {"pos": [[1, 1, 1], [22, 22, 22]], "m": [1, 1], "r": [1, 1], "G": "6.673e-11",
            "dist": dstr(dist), "F": dstr(F), "f0": [dstr(v) for v in f0],
            "f1": [dstr(-v) for v in f0], "v0_after_update": [dstr(v) for v in v0],
            "x0_after_update": [dstr(v) for v in x0]}


def elastic_closed(x1, x2, v1, v2, m1, m2, r12, R):
    """Impulse along the contact normal at first contact; restitution applied around the CM
    velocity like doElastic.  Returns (collides, v1', v2')."""
    u = vsub(v1, v2)
    x21 = vsub(x2, x1)
    uu = vdot(u, u)
    if uu == 0:
        return False, v1, v2
    # approaching: the component of u along x21 is positive (thetav < pi/2)
    if vdot(u, x21) <= 0:
        return False, v1, v2
    # |x21 - u t|^2 = r12^2 -> uu t^2 - 2 (u.x21) t + (x21.x21 - r12^2) = 0
    b = vdot(u, x21)
    c = vdot(x21, x21) - r12 * r12
    disc = b * b - uu * c
    if disc < 0:
        return False, v1, v2
    t = (b - disc.sqrt()) / uu
    n = [(x21[k] - u[k] * t) / r12 for k in range(3)]
    un = vdot(u, n)
    v1e = [v1[k] - 2 * m2 / (m1 + m2) * un * n[k] for k in range(3)]
    v2e = [v2[k] + 2 * m1 / (m1 + m2) * un * n[k] for k in range(3)]
    vcm = [(m1 * v1[k] + m2 * v2[k]) / (m1 + m2) for k in range(3)]
    v1r = [(v1e[k] - vcm[k]) * R + vcm[k] for k in range(3)]
    v2r = [(v2e[k] - vcm[k]) * R + vcm[k] for k in range(3)]
    return True, v1r, v2r


def case(x1, x2, v1, v2, m1, m2, r1, r2, R):
    Dl = lambda v: [D(repr(float(a))) for a in v]
    ok, a, b = elastic_closed(Dl(x1), Dl(x2), Dl(v1), Dl(v2), D(repr(float(m1))),
                              D(repr(float(m2))), D(repr(float(r1))) + D(repr(float(r2))),
                              D(repr(float(R))))
    return {"x1": list(map(float, x1)), "x2": list(map(float, x2)), "v1": list(map(float, v1)),
            "v2": list(map(float, v2)), "m1": float(m1), "m2": float(m2), "r1": float(r1),
            "r2": float(r2), "R": float(R), "collides": ok, "v1_after": [dstr(q) for q in a],
            "v2_after": [dstr(q) for q in b]}


def oblique_cases(seed=7, count=40):
    rng = random.Random(seed)
    out = []
    while len(out) < count:
        r1, r2 = rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0)
        r12 = r1 + r2
        # separation between 0.6 and 1.0 r12 (overlapping, as detection requires)
        dirn = [rng.gauss(0, 1) for _ in range(3)]
        nn = sum(q * q for q in dirn) ** 0.5
        d = rng.uniform(0.6, 1.0) * r12
        x1 = [rng.uniform(-5, 5) for _ in range(3)]
        x2 = [x1[k] + d * dirn[k] / nn for k in range(3)]
        v1 = [rng.uniform(-3, 3) for _ in range(3)]
        v2 = [rng.uniform(-3, 3) for _ in range(3)]
        m1, m2 = rng.uniform(0.2, 5), rng.uniform(0.2, 5)
        R = 1.0 if len(out) % 3 else rng.uniform(0.3, 1.0)
        c = case(x1, x2, v1, v2, m1, m2, r1, r2, R)
        # keep clear of grazing and perpendicular incidence, where the coll3d angles are
        # ill-conditioned
        u = [v1[k] - v2[k] for k in range(3)]
        x21 = [x2[k] - x1[k] for k in range(3)]
        uu = sum(q * q for q in u) ** 0.5
        cos_t = sum(u[k] * x21[k] for k in range(3)) / (uu * d)
        sin_t = max(0.0, 1 - cos_t * cos_t) ** 0.5
        if c["collides"] and d * sin_t / r12 < 0.95 and cos_t > 0.05:
            out.append(c)
    return out


def main():
    data = {
        "wp_compute": wp_compute(),
        "head_on": case([0, 0, 0], [2, 0, 0], [1, 0, 0], [0, 0, 0], 1, 1, 1, 1, 1),
        "one_d": case([0, 0, 0], [0, 0, 1.5], [0, 0, 2], [0, 0, -0.5], 1, 3, 1, 1, 1),
        "oblique": oblique_cases(),
        "collide": {"pos": [500, 500, 500], "note": "coincident bodies: NaN velocities, collided"},
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "known_answers.json")
    with open(path, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
