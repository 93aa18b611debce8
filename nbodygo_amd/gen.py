"""Seeded initial-condition generators for the benchmark and test workloads.

The reference's generators are clock-seeded (cmd/util/vectorutil.go:34, cmd/sim/simgen.go:114)
and so not reproducible; these are seeded (numpy PCG64) and fully determined by their arguments.

  uniform_sphere   positions uniform in a ball (rejection sampling as GetVectorEven,
                   vectorutil.go:32-42), lognormal masses, radii k * m^(1/3), Gaussian velocities.
  plummer          Plummer sphere (Aarseth, Henon & Wielen 1974, A&A 37, 183): radii from the
                   inverted cumulative mass profile, speeds by von Neumann rejection on
                   g(q) = q^2 (1 - q^2)^(7/2), isotropic directions, centre-of-mass frame.
  two_plummer      two Plummer spheres on a collision course.
  batch_clusters   independent uniform-sphere clusters of uneven size.
"""
import numpy as np

from . import _lib as L
from .sim import Bodies


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def _unit_vectors(rng, n):
    u = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    s = np.sqrt(1.0 - u * u)
    return np.stack([s * np.cos(phi), s * np.sin(phi), u], axis=1)


def _ball(rng, n, radius):
    out = np.empty((0, 3))
    while out.shape[0] < n:
        k = max(2 * (n - out.shape[0]), 16)
        p = rng.uniform(-radius, radius, (k, 3))
        p = p[np.sum(p * p, axis=1) <= radius * radius]
        out = np.concatenate([out, p])
    return out[:n]


def uniform_sphere(n, seed, radius=1.0e3, mass_mu=np.log(1.0e8), mass_sigma=1.0, radius_k=0.05,
                   vel_sigma=5.0, behavior=L.NB_CB_ELASTIC):
    """Uniform ball, lognormal masses, radii radius_k * m^(1/3), velocities N(0, vel_sigma^2)."""
    rng = _rng(seed)
    p = _ball(rng, n, radius)
    m = rng.lognormal(mass_mu, mass_sigma, n)
    r = radius_k * np.cbrt(m)
    v = rng.normal(0.0, vel_sigma, (n, 3))
    return Bodies(p[:, 0], p[:, 1], p[:, 2], v[:, 0], v[:, 1], v[:, 2], m, r,
                  beh=np.full(n, behavior, np.int32))


def plummer(n, seed, a=1.0, M=1.0, G=1.0, radius=0.0, mass_cut=0.999, center=(0, 0, 0),
            velocity=(0, 0, 0), behavior=L.NB_CB_ELASTIC):
    """Plummer sphere of scale a, total mass M (equal masses), in virial equilibrium for G.

    Radii come from r = a / sqrt(X^(-2/3) - 1) with X uniform in (0, mass_cut] (the cut drops the
    far tail, as Aarseth et al. do); every body gets the same collision radius `radius`."""
    rng = _rng(seed)
    X = rng.uniform(0.0, mass_cut, n)
    X = np.where(X <= 0.0, mass_cut, X)
    rr = a / np.sqrt(X ** (-2.0 / 3.0) - 1.0)
    pos = _unit_vectors(rng, n) * rr[:, None]
    # von Neumann rejection for q = v / v_esc: g(q) = q^2 (1 - q^2)^3.5 <= 0.1
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        x4 = rng.uniform(0.0, 1.0, todo.size)
        x5 = rng.uniform(0.0, 0.1, todo.size)
        ok = x5 < x4 * x4 * (1.0 - x4 * x4) ** 3.5
        q[todo[ok]] = x4[ok]
        todo = todo[~ok]
    vesc = np.sqrt(2.0 * G * M) * (rr * rr + a * a) ** -0.25
    vel = _unit_vectors(rng, n) * (q * vesc)[:, None]
    pos -= pos.mean(axis=0)
    vel -= vel.mean(axis=0)
    pos += np.asarray(center, np.float64)
    vel += np.asarray(velocity, np.float64)
    m = np.full(n, M / n)
    r = np.full(n, float(radius))
    return Bodies(pos[:, 0], pos[:, 1], pos[:, 2], vel[:, 0], vel[:, 1], vel[:, 2], m, r,
                  beh=np.full(n, behavior, np.int32))


def concat(*parts):
    keys = ("x", "y", "z", "vx", "vy", "vz", "m", "r", "beh", "flags", "rest", "frag", "fstep")
    return Bodies(*(np.concatenate([getattr(p, k) for p in parts]) for k in keys))


def two_plummer(n_each, seed, a=1.0, mass_ratio=3.0, offset=10.0, rel_velocity=0.5, G=1.0,
                radius_k=0.0):
    """Two Plummer spheres of n_each bodies, masses 1 and mass_ratio, centres `offset` apart on
    the x axis, approaching at rel_velocity, in the centre-of-mass frame.  Collision radii are
    radius_k * m^(1/3)."""
    m1, m2 = 1.0, float(mass_ratio)
    mt = m1 + m2
    c1 = np.array([-offset * m2 / mt, 0.0, 0.0])
    c2 = np.array([offset * m1 / mt, 0.0, 0.0])
    v1 = np.array([rel_velocity * m2 / mt, 0.0, 0.0])
    v2 = np.array([-rel_velocity * m1 / mt, 0.0, 0.0])
    a1 = plummer(n_each, seed, a=a, M=m1, G=G, center=c1, velocity=v1)
    a2 = plummer(n_each, seed + 1, a=a * (m2 / m1) ** (1.0 / 3.0), M=m2, G=G, center=c2,
                 velocity=v2)
    b = concat(a1, a2)
    b.r[:] = radius_k * np.cbrt(b.m)
    return b


def batch_sizes(nsims, seed, lo=256, hi=8192):
    """Sim sizes drawn uniform in [lo, hi]."""
    return _rng(seed).integers(lo, hi + 1, nsims)


def batch_clusters(sizes, seed, radius=1.0, mass_mu=0.0, mass_sigma=1.0, vel_sigma=0.1,
                   radius_k=0.0):
    """One uniform-sphere cluster with lognormal masses per entry of sizes."""
    return [uniform_sphere(int(n), seed * 100003 + k, radius=radius, mass_mu=mass_mu,
                           mass_sigma=mass_sigma, radius_k=radius_k, vel_sigma=vel_sigma)
            for k, n in enumerate(sizes)]
