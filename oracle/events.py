"""CPU restatement of the reference's subsume and fragmentation events -- TEST INFRASTRUCTURE ONLY.

One cycle in the runner's phase order (computation-runner.go:267-326) with the event semantics
of the device path's nb_params.events = 1, over the C oracle's force, coll3d (or_elastic) and
update:

  Compute     a fragmenting body runs fragment() (fragcalc.go:97-117) and nothing else; every
              other active body gets its force and, for each overlapping partner, Compute's
              dispatch (body.go:171-186): a collision event when both are Elastic or Fragment,
              else a subsume event when either is Subsume and the larger body (radii unequal)
              reaches the other's centre
  ProcessMods each event pair once, in ascending (i, j) order: ResolveSubsume (body.go:228-244,
              no existence check) or ResolveCollision (body.go:248-264) whose shouldFragment
              (fragcalc.go:24-49) starts fragmentation (doFragment, initiateFragmentation:
              fragcalc.go:54-83) instead of doElastic
  Update      Body.Update (body.go:114-139); a fragmenting body is kicked by the force of the
              cycle it started fragmenting in (the reference's Body.fx is not reset for it)
  Cycle       compaction, then the fragments newest first, then queued adds newest first
              (body_collection.go:253-296)

Constants as the reference has them: at most 2000 fragments, 101 created per cycle, fragment
radius 1 (Pow(x, 1/3) with the integer 1/3 == 0), FragFactor > 10 caps fragDelta at 10.
Fragment positions: GetVectorEven's rejection draw (vectorutil.go:32-42) from the same
counter-based generator as the device (the reference seeds math/rand from the clock).

Pinned by hand-worked subsume chains and fragment counts (tests/test_events_oracle.py), by the
generator's uniformity and ball containment, and by mass accounting.
"""
import math

import numpy as np

from . import binding as B

M64 = (1 << 64) - 1
EXISTS, FRAGMENTING, SUN, PINNED = 1, 2, 4, 8
NONE, SUBSUME, ELASTIC, FRAGMENT = 0, 1, 2, 3
FRAG_PER_CYCLE = 101
FRAG_RADIUS = 1.0
INT64_MIN = -(1 << 63)


def mix(z):
    """SplitMix64 finaliser."""
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def uniform(fid, ordinal, draw):
    """Uniform double in [0, 1) for (fragmentation id, fragment ordinal, draw index)."""
    h = mix(mix(fid) ^ ((ordinal * 0xD1B54A32D192ED03) & M64) ^
            ((draw * 0xA0761D6478BD642F) & M64))
    return float(h >> 11) * 2.0 ** -53


def vector_even(fid, ordinal):
    """GetVectorEven's unit-ball draw (vectorutil.go:35-40): x, y, z in [-1, 1) until
    x^2 + y^2 + z^2 <= 1 (at most 64 rounds)."""
    x = y = z = 0.0
    for d in range(64):
        x = uniform(fid, ordinal, 3 * d) * 2 - 1
        y = uniform(fid, ordinal, 3 * d + 1) * 2 - 1
        z = uniform(fid, ordinal, 3 * d + 2) * 2 - 1
        if not (x * x + y * y + z * z > 1):
            break
    return x, y, z


def can_collide(b):
    return b == ELASTIC or b == FRAGMENT


def pair_event(bi, bj, ri, rj, dist):
    if can_collide(bi) and can_collide(bj):
        return True
    if bi != SUBSUME and bj != SUBSUME:
        return False
    return (ri > rj and dist <= ri) or (rj > ri and dist <= rj)


FIELDS = ("x", "y", "z", "vx", "vy", "vz", "m", "r", "rest", "frag")


def ieee_div(a, b):
    """a / b with IEEE results for a zero divisor (inf, or NaN for 0/0), as Go and the device."""
    with np.errstate(divide="ignore", invalid="ignore"):
        return float(np.float64(a) / np.float64(b))


class EventSim:
    """Bodies as numpy arrays (x..frag, fstep, beh, flags) plus fragmentation state."""

    def __init__(self, b, G, eps=0.0, cutoff=True):
        self.G, self.eps, self.cutoff = G, eps, cutoff
        self._bind({k: np.array(getattr(b, k), np.float64) for k in FIELDS},
                   np.array(b.beh, np.int32), np.array(b.flags, np.uint8))
        self.fstep = np.array(b.fstep, np.float64)
        n = len(self.beh)
        self.fs_f = np.zeros((n, 3))      # kept force
        self.fs_mass = np.zeros(n)
        self.fs_radius = np.zeros(n)
        self.fs_pos = np.zeros((n, 3))
        self.fs_left = np.zeros(n, np.int64)
        self.fs_made = np.zeros(n, np.int64)
        self.fs_id = np.zeros(n, np.int64)
        self.seq = 0
        self.pending = []                 # fragments in creation order
        self.stats = dict(subsumed=0, fragmentations=0, emitted=0, resolved=0, requests=0)

    @property
    def n(self):
        return len(self.beh)

    def active(self):
        return ((self.flags & EXISTS) != 0) & ((self.flags & FRAGMENTING) == 0)

    def _bind(self, a, beh, flags):
        """The C oracle's state over these arrays (its own copies, which become the store)."""
        self.st = B.State(a["x"], a["y"], a["z"], a["vx"], a["vy"], a["vz"], a["m"], a["r"],
                          beh=beh, flags=flags, rest=a["rest"], frag=a["frag"])
        self.a = self.st.a
        self.beh = self.st.beh
        self.flags = self.st.flags

    # -- phases ------------------------------------------------------------------------------
    def emit(self):
        """fragment() of every fragmenting body, in body order (fragcalc.go:97-117)."""
        a = self.a
        for i in range(self.n):
            f = self.flags[i]
            if not ((f & EXISTS) and (f & FRAGMENTING)):
                continue
            left = int(self.fs_left[i])
            k = min(left, FRAG_PER_CYCLE) if left > 0 else 0
            rad = self.fs_radius[i] * .9
            for q in range(k):
                x, y, z = vector_even(int(self.fs_id[i]), int(self.fs_made[i]) + q)
                self.pending.append((
                    (x * rad) + self.fs_pos[i, 0], (y * rad) + self.fs_pos[i, 1],
                    (z * rad) + self.fs_pos[i, 2], a["vx"][i], a["vy"][i], a["vz"][i],
                    self.fs_mass[i], FRAG_RADIUS))
            self.fs_left[i] = left - k
            self.fs_made[i] += k
            self.stats["emitted"] += k
            if left - k <= 0:
                self.flags[i] = f & np.uint8(~EXISTS & 0xFF)

    def force(self):
        st = self.st
        n = self.n
        f = np.zeros((3, n))
        fp = [f[c].ctypes.data_as(B._dp) for c in range(3)]
        import ctypes as C
        if self.eps == 0.0 and self.cutoff:
            B.lib().or_force_ref(C.byref(st.s), self.G, None, 0, *fp, 16)
        else:
            B.lib().or_force_soft(C.byref(st.s), self.G, self.eps, int(self.cutoff), None, 0,
                                  *fp, 16)
        return f.T.copy()

    def detect(self):
        """Event pairs, canonical (i < j)."""
        a = self.a
        act = self.active()
        idx = np.nonzero(act)[0]
        out = []
        x, y, z, r = a["x"], a["y"], a["z"], a["r"]
        for i in idx:
            js = idx[idx > i]
            if js.size == 0:
                continue
            dx = x[js] - x[i]
            dy = y[js] - y[i]
            dz = z[js] - z[i]
            dist = np.sqrt(dx * dx + dy * dy + dz * dz)
            over = ~(dist > r[i] + r[js])
            for j, d in zip(js[over], dist[over]):
                if pair_event(int(self.beh[i]), int(self.beh[j]), r[i], r[j], d):
                    out.append((int(i), int(j)))
        return np.array(out, np.int32).reshape(-1, 2)

    def _initiate(self, b, factor, force):
        a = self.a
        frag_delta = 10.0 if a["frag"][b] > 10 else factor - a["frag"][b]
        prod = frag_delta * self.fstep[b]
        fragments = prod if math.isnan(prod) else min(prod, 2000.0)
        if fragments <= 1:
            self.beh[b] = FRAGMENT
            return
        if not (self.flags[b] & FRAGMENTING):
            self.stats["fragmentations"] += 1
        self.flags[b] |= FRAGMENTING
        self.fs_f[b] = force[b]
        self.fs_mass[b] = a["m"][b] / fragments
        self.fs_radius[b] = a["r"][b]
        self.fs_pos[b] = (a["x"][b], a["y"][b], a["z"][b])
        self.fs_left[b] = INT64_MIN if math.isnan(fragments) else int(fragments)
        self.fs_made[b] = 0
        self.fs_id[b] = self.seq
        self.seq += 1

    def resolve(self, pairs, force, collided):
        a = self.a
        for lo, hi in pairs:
            if not (can_collide(self.beh[lo]) and can_collide(self.beh[hi])):
                b, o = (lo, hi) if a["r"][lo] > a["r"][hi] else (hi, lo)
                a["m"][b] = a["m"][b] + a["m"][o]
                self.flags[o] &= np.uint8(~EXISTS & 0xFF)
                a["m"][o] = 0.0
                self.stats["subsumed"] += 1
                continue
            if self.beh[lo] == ELASTIC and can_collide(self.beh[hi]):
                b, o = lo, hi
            elif self.beh[hi] == ELASTIC and can_collide(self.beh[lo]):
                b, o = hi, lo
            else:
                continue
            if not (self.flags[b] & EXISTS) or not (self.flags[o] & EXISTS):
                continue
            ok, r = B.elastic(self.st, int(b), int(o))
            if not ok:
                continue
            self.stats["resolved"] += 1
            frag_b, frag_o = self.beh[b] == FRAGMENT, self.beh[o] == FRAGMENT
            if frag_b or frag_o:  # shouldFragment (fragcalc.go:24-49)
                br = a["rest"][b]
                vt = a["vx"][b] + a["vy"][b] + a["vz"][b]
                dvt = (abs(a["vx"][b] - ((r[0] - r[6]) * br + r[6])) +
                       abs(a["vy"][b] - ((r[1] - r[7]) * br + r[7])) +
                       abs(a["vz"][b] - ((r[2] - r[8]) * br + r[8])))
                tf = ieee_div(dvt, abs(vt))
                vo = a["vx"][o] + a["vy"][o] + a["vz"][o]
                dvo = (abs(a["vx"][o] - ((r[3] - r[6]) * br + r[6])) +
                       abs(a["vy"][o] - ((r[4] - r[7]) * br + r[7])) +
                       abs(a["vz"][o] - ((r[5] - r[8]) * br + r[8])))
                of = ieee_div(dvo, abs(vo))
                if (frag_b and tf > a["frag"][b]) or (frag_o and of > a["frag"][o]):
                    self.stats["requests"] += 1
                    if frag_b and tf > a["frag"][b]:
                        self._initiate(b, tf, force)
                    if frag_o and of > a["frag"][o]:
                        self._initiate(o, of, force)
                    continue
            rr = a["rest"][b]
            a["vx"][b] = (r[0] - r[6]) * rr + r[6]
            a["vy"][b] = (r[1] - r[7]) * rr + r[7]
            a["vz"][b] = (r[2] - r[8]) * rr + r[8]
            a["vx"][o] = (r[3] - r[6]) * rr + r[6]
            a["vy"][o] = (r[4] - r[7]) * rr + r[7]
            a["vz"][o] = (r[5] - r[8]) * rr + r[8]
            collided[b] = collided[o] = 1

    def update(self, force, collided, dt, R):
        f = force.copy()
        frag = (self.flags & FRAGMENTING) != 0
        f[frag] = self.fs_f[frag]
        self.st.collided[:] = collided
        B.update(self.st, f, dt, R)

    def step(self, dt, R):
        """One cycle; returns the event pairs."""
        self.emit()
        force = self.force()
        pairs = self.detect()
        collided = np.zeros(self.n, np.uint8)
        self.resolve(pairs, force, collided)
        self.update(force, collided, dt, R)
        return pairs

    def cycle(self, R):
        """Compaction, then the fragments newest first (rest = R)."""
        keep = (self.flags & EXISTS) != 0
        fr = self.pending[::-1]
        k = len(fr)
        cat = lambda arr, tail: np.concatenate([arr[keep], tail])
        cols = list(zip(*fr)) if k else [[]] * 8
        a = {}
        for c, key in enumerate(("x", "y", "z", "vx", "vy", "vz", "m", "r")):
            a[key] = cat(self.a[key], np.array(cols[c], np.float64))
        a["rest"] = cat(self.a["rest"], np.full(k, R))
        a["frag"] = cat(self.a["frag"], np.zeros(k))
        self._bind(a, cat(self.beh, np.full(k, ELASTIC, np.int32)).astype(np.int32),
                   cat(self.flags, np.full(k, EXISTS, np.uint8)).astype(np.uint8))
        self.fstep = cat(self.fstep, np.zeros(k))
        self.fs_f = np.concatenate([self.fs_f[keep], np.zeros((k, 3))])
        self.fs_pos = np.concatenate([self.fs_pos[keep], np.zeros((k, 3))])
        for name in ("fs_mass", "fs_radius"):
            setattr(self, name, cat(getattr(self, name), np.zeros(k)))
        for name in ("fs_left", "fs_made", "fs_id"):
            setattr(self, name, cat(getattr(self, name), np.zeros(k, np.int64)))
        self.pending = []
        return self.n
