"""Benchmark of the per-cycle compute path on one MI355X.

Default workload: a Plummer sphere of N = 262,144 bodies (a = 1, M = 1, G = 1), eps = 0.01,
dt = 1e-3, collisions off.  A "step" is one full cycle: force (+ detection), resolution, update.
Inputs are resident in HBM before timing.  The line compares, on that sphere:
  - the plain step: FP32 pairwise force with Kahan tile sums, FP64 integration, in the kernel
    shape the autotuner picks (the headline `value`: G interactions/s, ordered pairs N(N-1) per
    step, the reference's nbody_computations counter; `roofline` is this force kernel's FP32
    fraction, its mean launch time taken with HIP events on the library's stream);
  - the learned far-field surrogate step (`surrogate`: 32^3 stretched grid, exact near field,
    MLP residual trained on direct-sum rows of 200 seeded sims of the same size, sampled error
    check and fallback inside the timed steps), with its force error against FP64 as the global
    norm, the per-body p99 and the maximum.
Beside it, measured in the same run: the 65,536-body sphere with collisions off and elastic
(`plummer65k`), 512 batched sims (`batch512`), a short run of two colliding spheres of 524,288
bodies each (`collide2x524k`), and the reference's CPU algorithm (its worker pool of threads
over contiguous target slices, restated) on a bounded sample of the headline workload and on
the 1,000-body FP64 scene for 100 full cycles (`cpu_baseline`, `cpu_config1`), and an optimised
vectorised FP64 direct sum as a stronger CPU line (`cpu_soa`).

Other workloads (`--workload`) print one line of their own.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "G interactions/s + steps/s, plain vs surrogate at ≤1e-3 rel force err"
PEAK_FP32 = 157.3e12  # MI355X FP32 vector peak (FLOP/s)
PEAK_FP64 = 78.6e12   # MI355X FP64 vector peak (FLOP/s)
# "mfma" is the output line's label for a compute-bound kernel; what bounds these kernels is the
# vector ALU, whose published FP32 and FP64 peaks equal the matrix cores' (157.3, 78.6 TFLOP/s).
BOUND_FP32 = "compute: FP32 vector ALU (v_fma_f32, v_rsq_f32), peak 157.3 TFLOP/s"
BOUND_FP64 = "compute: FP64 vector ALU (divisions, square roots), peak 78.6 TFLOP/s"
FLOPS_PER_INTERACTION = 20
ROOT = os.path.dirname(os.path.abspath(__file__))
_T0 = time.perf_counter()


def log(msg):
    """Progress on stderr (the JSON line alone goes to stdout)."""
    print("bench [%6.1f s] %s" % (time.perf_counter() - _T0, msg), file=sys.stderr, flush=True)

WORKLOADS = {
    "plummer262k": dict(n=262144, desc="Plummer sphere N=262144 (a=1, M=1, G=1), eps=0.01, "
                                       "dt=1e-3, collisions off: plain FP32 direct sum (Kahan "
                                       "tiles, FP64 update) against the learned far-field "
                                       "surrogate (32^3 stretched grid, exact near field, MLP "
                                       "residual, sampled check every step)"),
    "plummer65k": dict(n=65536, desc="Plummer sphere N=65536 (a=1, M=1, G=1), eps=0.01, dt=1e-3, "
                                     "FP32 force + Kahan, FP64 update; collisions off, then "
                                     "elastic at radius 1e-4 and at radius 5e-3"),
    "surrogate262k": dict(n=262144, desc="Plummer sphere N=262144 (a=1, M=1, G=1), eps=0.01, "
                                         "dt=1e-3, collisions off: learned far-field surrogate "
                                         "(32^3 stretched grid, exact near field, MLP residual, "
                                         "sampled check every step) against the plain step"),
    "collide2x524k": dict(n=1048576, desc="two Plummer spheres of 524,288 bodies (masses 1:3, "
                                          "offset 10a, approaching at 0.5), elastic collisions, "
                                          "radii 0.05 m^(1/3), eps=0.01, dt=1e-3, 64^3 grid, "
                                          "shell 3: surrogate far field + exact near field "
                                          "(detection there) against the plain FP32 direct sum; "
                                          "energy and momentum drift from FP64 diagnostics"),
    "barneshut262k": dict(n=262144, desc="Plummer sphere N=262144 (a=1, M=1, G=1), dt=1e-3, "
                                         "collisions off: the reference's Barnes-Hut octree "
                                         "(theta 0.3, softening 1e-4, Build wired in), exact "
                                         "FP64 walk and fast group walk, against the FP64 "
                                         "direct sum with the same softening"),
    "lifecycle1m": dict(n=1048576, desc="Plummer sphere N=1,048,576, FP32 path: one delete "
                                       "request of 1 % of the bodies (processDeletes' walk), "
                                       "1,000 queued adds and Cycle, on the device, against "
                                       "the same through the host (read back, compact, "
                                       "reload)"),
    "dense_collisions": dict(n=20000, desc="dense elastic scenes of the 1,000-body kind (uniform "
                                           "sphere of radius 1e3, lognormal masses) at larger N "
                                           "and radii, one cycle each: the collision pass along "
                                           "the pairs' dependency DAG against one pair at a "
                                           "time (HIP-event time of the resolution kernels)"),
    "batch512": dict(n=0, desc="512 independent simulations, N uniform in [256, 8192] (seeded), "
                               "uniform-sphere clusters with lognormal masses, eps=0.01, "
                               "dt=1e-3, FP32 force + Kahan, FP64 update, one hipGraph launch "
                               "per call, autotuned tile shapes"),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="plummer262k", choices=sorted(WORKLOADS))
    ap.add_argument("--headline-only", action="store_true",
                    help="time the headline plain step alone (no surrogate, extras or CPU legs)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-tune", action="store_true", help="use the default kernel shape")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--shape", default="",
                    help="pin the FP32 force kernel shape: ipl,tile_j,jsplit,block,variant")
    ap.add_argument("--no-elastic", action="store_true", help="skip the elastic-collision runs")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate comparison")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the plummer65k, batch512 and collide2x524k legs")
    ap.add_argument("--sur-sims", type=int, default=200, help="training sims for the surrogate")
    ap.add_argument("--sur-steps", type=int, default=50, help="timed surrogate steps")
    ap.add_argument("--sur-fit-steps", type=int, default=4000, help="Adam steps of the MLP fit")
    ap.add_argument("--c5-steps", type=int, default=0,
                    help="cycles of the collision run (0: 200 alone, 20 inside the default line)")
    ap.add_argument("--c5-tol", type=float, default=1e-3,
                    help="sampled-error tolerance of the surrogate in the collision run")
    ap.add_argument("--c5-train-sims", type=int, default=8, help="training sims, collision run")
    ap.add_argument("--c5-grid", type=int, default=64, help="surrogate cells per axis, collision run")
    ap.add_argument("--c5-shell", type=int, default=3, help="surrogate shell depth, collision run")
    return ap.parse_args()


def shape_kw(spec):
    if not spec:
        return {}
    v = [int(q) for q in spec.split(",")]
    return dict(zip(("ipl", "tile_j", "jsplit", "block", "variant"), v))


def error_summary(a, ref):
    """Global ||da||/||a|| and per-body relative percentiles of a against ref (n x 3)."""
    d = np.linalg.norm(a - ref, axis=1)
    per = d / np.linalg.norm(ref, axis=1)
    return {"global": float(np.linalg.norm(a - ref) / np.linalg.norm(ref)),
            "p50": float(np.percentile(per, 50)), "p99": float(np.percentile(per, 99)),
            "max": float(per.max()), "targets": int(a.shape[0])}


def force_error(g, b, nsample=4096, seed=0):
    """FP32-path force error against the device FP64 direct sum (which the tests pin to the CPU
    oracle) on sampled targets: global ||da||/||a|| and per-body percentiles."""
    idx = np.sort(np.random.default_rng(seed).choice(b.n, min(nsample, b.n), replace=False))
    a32 = g.forces()[idx] / b.m[idx, None]
    a64 = g.accel_fp64(idx)
    out = error_summary(a32, a64)
    out["against"] = "FP64 direct sum on the device (bit-identical to the C oracle in tests)"
    return out


def measured_traffic(n, shape):
    """HBM bytes per force launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    (calibrated against a copy kernel of known size), when they were taken at this body count
    and kernel shape; None otherwise."""
    name = "force_plummer%dk_traffic.json" % (n // 1000)
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None, None
    if t.get("n") != n:
        return None, "profiles/" + name + " (taken at another body count: null)"
    for e in t.get("shapes", [t]):  # one entry per measured shape (or a single measurement)
        if e.get("shape") == shape:
            return e["bytes_per_launch"], "profiles/" + name
    return None, "profiles/" + name + " (taken at other shapes: null)"


def timed_run(g, b, steps, warmup, dt=1e-3):
    g.load(b)
    if warmup:
        g.step(dt, 1.0, nsteps=warmup)
    g.set_timing(True)
    t0 = time.perf_counter()
    s = g.step(dt, 1.0, nsteps=steps)
    t1 = time.perf_counter()
    g.set_timing(False)
    return s, t1 - t0


def plain_leg(n, steps, warmup, pinned, tune, seed=2024):
    """The plain FP32 step on a Plummer sphere of n bodies: rate, roofline of the force kernel,
    force error against FP64."""
    from nbodygo_amd import NBodyGPU, gen
    b = gen.plummer(n, seed=seed)
    g = NBodyGPU(G=1.0, eps=0.01, overlap_cutoff=False, collisions=False, **pinned)
    g.load(b)
    shape = g.autotune(reps=2) if (tune and not pinned) else g.config()
    s, el = timed_run(g, b, steps, warmup)
    ms_step = el / steps * 1e3
    launch_ms = s["force_ms"] / max(s["force_launches"], 1)
    per_launch = n * (n - 1)
    achieved = FLOPS_PER_INTERACTION * per_launch / (launch_ms * 1e-3)
    traffic, tsrc = measured_traffic(n, shape)
    out = {
        "n": n, "value": s["interactions"] / el / 1e9, "unit": "G interactions/s",
        "ms_per_step": ms_step, "steps_per_s": 1e3 / ms_step, "kernel_shape": shape,
        "roofline": {"bound": "mfma", "bound_detail": BOUND_FP32, "achieved": achieved / 1e12,
                     "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32,
                     "traffic": traffic, "traffic_source": tsrc,
                     "kernel": "k_force_lds / k_force_scalar (FP32 pairwise force)",
                     "flops_per_interaction": FLOPS_PER_INTERACTION,
                     "interactions_per_launch": per_launch, "avg_launch_ms": launch_ms},
        "force_error": force_error(g, b),
    }
    g.close()
    return out


def elastic_leg(n, radius, steps, warmup, pinned, tune):
    """The same sphere with elastic collisions (overlap cutoff on) at a given body radius:
    detected pairs and FP64 fix-up tiles per step next to the rate."""
    from nbodygo_amd import NBodyGPU, gen
    ge = NBodyGPU(G=1.0, eps=0.01, overlap_cutoff=True, collisions=True, **pinned)
    be = gen.plummer(n, seed=2024, radius=radius)
    ge.load(be)
    if tune and not pinned:
        ge.autotune(reps=2)
    se, ele = timed_run(ge, be, steps, warmup)
    ge.close()
    return {"value": se["interactions"] / ele / 1e9, "unit": "G interactions/s",
            "ms_per_step": ele / steps * 1e3, "radius": radius,
            "pairs_per_step": se["pairs_total"] / steps,
            "resolved_per_step": se["resolved_total"] / steps,
            "fixup_tiles_per_step": se["fixup_tiles"] / steps}


def surrogate_leg(a, pinned, plain=None):
    """Plain vs surrogate at N = 262,144 (Plummer, eps = 0.01, dt = 1e-3, collisions off)."""
    from nbodygo_amd import NBodyGPU, gen
    n = 262144
    b = gen.plummer(n, seed=2024)
    g = NBodyGPU(G=1.0, eps=0.01, overlap_cutoff=False, collisions=False, **pinned)
    if plain is None:
        g.load(b)
        shape = g.config() if (a.no_tune or pinned) else g.autotune(reps=2)
        ps, pel = timed_run(g, b, 10, 3)
        plain = {"ms_per_step": pel / 10 * 1e3, "value": ps["interactions"] / pel / 1e9,
                 "kernel_shape": shape}
    # training rows: 2048 bodies from each of the seeded sims, labelled by the FP32 direct sum
    t0 = time.perf_counter()
    for k in range(a.sur_sims):
        g.load(gen.plummer(n, seed=100000 + k))
        g.sur_enable(check_every=0)
        g.sur_collect(2048, seed=k)
    t_collect = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss = g.sur_fit(steps=a.sur_fit_steps, batch=512, lr=1e-3)
    t_fit = time.perf_counter() - t0
    rows = g.sur_stats()["train_rows"]
    # held-out state: error without / with the learned residual, against FP64
    g.load(b)
    g.sur_enable(check_every=1, tol=1e-3)
    err_nomlp = g.sur_error(4096, use_mlp=False)
    err_mlp = force_error(g, b)
    s0 = g.sur_stats()
    ss, sel = timed_run(g, b, a.sur_steps, 3)
    s1 = g.sur_stats()
    sur_ms = sel / a.sur_steps * 1e3
    g.close()
    return {
        "n": n, "workload": "Plummer sphere N=262144 (a=1, M=1, G=1), eps=0.01, dt=1e-3, "
                            "collisions off; 32^3 stretched grid, shell 2, MLP 3x128 GELU",
        "ms_per_step": sur_ms, "steps_per_s": 1e3 / sur_ms,
        "equiv_G_interactions_per_s": ss["interactions"] / sel / 1e9,
        "plain_ms_per_step": plain["ms_per_step"], "plain_steps_per_s": 1e3 / plain["ms_per_step"],
        "plain_G_interactions_per_s": plain["value"],
        "plain_kernel_shape": plain["kernel_shape"],
        "speedup_vs_plain": plain["ms_per_step"] / sur_ms,
        "force_error": err_mlp, "error_without_mlp": err_nomlp,
        "checks": s1["checks"] - s0["checks"], "fallbacks": s1["fallbacks"] - s0["fallbacks"],
        "tolerance": 1e-3, "timed_steps": a.sur_steps,
        "training": {"sims": a.sur_sims, "rows": rows, "adam_steps": a.sur_fit_steps, "batch": 512,
                     "loss": loss, "collect_s": t_collect, "fit_s": t_fit},
    }


def batch_leg(steps, warmup, tune=True):
    """Config 3: the batched path; a step advances all 512 simulations by one cycle."""
    from nbodygo_amd import NBodyBatch, gen
    sizes = gen.batch_sizes(512, seed=2024)
    sims = gen.batch_clusters(sizes, seed=2024)
    bt = NBodyBatch(G=1.0, eps=0.01)
    bt.load(sims)
    shapes = bt.autotune(reps=2) if tune else bt.shapes()
    if warmup:
        bt.step(1e-3, 1.0, nsteps=warmup)
    t0 = time.perf_counter()
    s = bt.step(1e-3, 1.0, nsteps=steps)
    el = time.perf_counter() - t0
    bt.close()
    per_step = s["interactions"] / steps
    ms_step = el / steps * 1e3
    achieved = FLOPS_PER_INTERACTION * per_step / (ms_step * 1e-3)
    return {"value": s["interactions"] / el / 1e9, "unit": "G interactions/s",
            "ms_per_step": ms_step, "steps_per_s": 1e3 / ms_step, "sims": 512,
            "bodies": int(sizes.sum()), "shapes": shapes,
            "frac_of_fp32_peak_whole_step": achieved / PEAK_FP32}


def collide_leg(a, pinned, steps):
    """Config 5: two colliding Plummer spheres, `steps` cycles with the surrogate (collisions
    detected in its exact near field) and the same with the plain direct sum, from the same
    start."""
    from nbodygo_amd import NBodyGPU, gen
    n_each, dt, tol = 524288, 1e-3, a.c5_tol
    b = gen.two_plummer(n_each, seed=2024, radius_k=0.05)
    pscale = float(np.sum(b.m * np.linalg.norm(b.vel(), axis=1)))

    def drift(d0, d1):
        e0 = d0["kinetic"] + d0["potential"]
        e1 = d1["kinetic"] + d1["potential"]
        return {"energy_rel": abs(e1 - e0) / abs(e0),
                "momentum_rel": float(np.linalg.norm(d1["momentum"] - d0["momentum"]) / pscale),
                "E0": e0, "E1": e1}

    # the plain path runs the LDS-staged kernel shape the autotuner picks at large N
    g = NBodyGPU(G=1.0, eps=0.01, overlap_cutoff=False, collisions=True,
                 **(pinned or dict(ipl=1, tile_j=512, jsplit=1, block=256, variant=2)))
    t0 = time.perf_counter()
    for k in range(a.c5_train_sims):
        g.load(gen.two_plummer(n_each, seed=300000 + k, radius_k=0.05))
        g.sur_enable(check_every=0, grid=a.c5_grid, shell=a.c5_shell)
        g.sur_collect(4096, seed=k)
    loss = g.sur_fit(steps=a.sur_fit_steps, batch=512, lr=1e-3) if a.c5_train_sims else None
    t_train = time.perf_counter() - t0
    out = {}
    for mode in ("surrogate", "plain"):
        g.load(b)
        if mode == "surrogate":
            g.sur_enable(check_every=1, tol=tol, grid=a.c5_grid, shell=a.c5_shell)
            err_mlp = force_error(g, b, nsample=1024)
            err_base = g.sur_error(1024, use_mlp=False)
        else:
            g.sur_disable()
            err_mlp = force_error(g, b, nsample=1024)
            err_base = None
        d0 = g.diagnostics()
        s0 = g.sur_stats()
        t0 = time.perf_counter()
        st = g.step(dt, 1.0, nsteps=steps)
        el = time.perf_counter() - t0
        d1 = g.diagnostics()
        s1 = g.sur_stats()
        out[mode] = {"ms_per_step": el / steps * 1e3, "steps_per_s": steps / el,
                     "force_error_start": err_mlp, "pairs_total": st["pairs_total"],
                     "resolved_total": st["resolved_total"], "drift": drift(d0, d1),
                     "flagged_chunks_per_step": st["fixup_tiles"] / steps,
                     "final": g.store()}
        if mode == "surrogate":
            out[mode]["error_without_mlp"] = err_base
            out[mode]["checks"] = s1["checks"] - s0["checks"]
            out[mode]["fallbacks"] = s1["fallbacks"] - s0["fallbacks"]
            # exact near-field pairs per cycle (the near kernel's work; the checks add none)
            out[mode]["near_pairs_per_step"] = (s1["near_pairs_total"] - s0["near_pairs_total"]) / steps
    g.close()
    fs, fp = out["surrogate"].pop("final"), out["plain"].pop("final")
    dpos = float(np.sqrt(np.mean(np.sum((fs.pos() - fp.pos()) ** 2, axis=1))))
    sur, pl = out["surrogate"], out["plain"]
    return {"n": 2 * n_each, "steps": steps, "tolerance": tol, "surrogate_grid": a.c5_grid,
            "surrogate_shell": a.c5_shell, "surrogate": sur, "plain": pl,
            "speedup_vs_plain": pl["ms_per_step"] / sur["ms_per_step"],
            "rms_position_difference_after": dpos,
            "training": {"sims": a.c5_train_sims, "rows_per_sim": 4096, "loss": loss,
                         "seconds": t_train}}


def run_collide(a, w, pinned):
    steps = a.c5_steps or 200
    c = collide_leg(a, pinned, steps)
    sur = c["surrogate"]
    print(json.dumps({
        "metric": METRIC, "value": 1e3 / sur["ms_per_step"], "unit": "steps/s (surrogate)",
        "n_gpus": 1, "steps": steps, "warmup": 0, "ms_per_step": sur["ms_per_step"],
        "steps_per_s": sur["steps_per_s"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 force (near field, expansions), bf16 MFMA (MLP), "
                                        "fp64 update and collisions",
        "data": "synthetic (seeded Plummer spheres)",
        "config": {"workload": w["desc"], "n": c["n"], "tolerance": c["tolerance"],
                   "surrogate_grid": a.c5_grid, "surrogate_shell": a.c5_shell},
        **{k: c[k] for k in ("surrogate", "plain", "speedup_vs_plain",
                             "rms_position_difference_after", "training")}}))


FLOPS_PER_BH_TERM = 23   # 3 sub, 3 mul + 3 add (softened r^2), sqrt, 2h and /dist (opening
                         # test), G*m*M and /r^2, 3 x (*d, /dist), 3 add; sqrt and / count as 1


def run_barneshut(a, w):
    """The reference's own far-field design: the octree with Build wired in, FP64, against the
    FP64 direct sum with the same softening at the same N; exact mode (the reference's walk and
    arithmetic per body) and fast mode (one walk per 64-body group, FP32 terms)."""
    from nbodygo_amd import NBodyGPU, gen
    n = w["n"]
    b = gen.plummer(n, seed=2024)
    idx = np.sort(np.random.default_rng(0).choice(n, 2048, replace=False))
    modes = {}
    for mode in ("exact", "fast"):
        g = NBodyGPU(G=1.0, eps=1e-4, overlap_cutoff=False, precision="fp64", collisions=False)
        g.load(b)
        g.bh_enable(mode=mode)
        s, el = timed_run(g, b, a.steps, a.warmup)
        st = g.bh_stats()
        ms_step = el / a.steps * 1e3
        launch_ms = s["force_ms"] / max(s["force_launches"], 1)
        g.load(b)
        f = g.forces()[idx] / b.m[idx, None]
        ref = g.accel_fp64(idx)
        g.close()
        modes[mode] = {"ms_per_step": ms_step, "force_ms": launch_ms,
                       "terms_per_step": st["interactions"], "terms_per_body": st["interactions"] / n,
                       "G_terms_per_s": st["interactions"] / (launch_ms * 1e-3) / 1e9,
                       "force_error_vs_fp64_direct": error_summary(f, ref),
                       "cells": st["nodes"], "depth_max": st["depth_max"]}
        log("barneshut %s: %.3f ms/step" % (mode, ms_step))
    h = NBodyGPU(G=1.0, eps=1e-4, overlap_cutoff=False, precision="fp64", collisions=False)
    sd, eld = timed_run(h, b, 2, 1)
    h.close()
    direct_ms = eld / 2 * 1e3
    ex, fa = modes["exact"], modes["fast"]
    achieved = FLOPS_PER_BH_TERM * ex["terms_per_step"] / (ex["force_ms"] * 1e-3)
    for m in modes.values():
        m["speedup_vs_fp64_direct"] = direct_ms / m["ms_per_step"]
    fa["speedup_vs_exact"] = ex["ms_per_step"] / fa["ms_per_step"]
    print(json.dumps({
        "metric": METRIC, "value": n * (n - 1) / (ex["ms_per_step"] * 1e-3) / 1e9,
        "unit": "G interactions/s (N(N-1) per step, Barnes-Hut exact mode)", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": ex["ms_per_step"],
        "steps_per_s": 1e3 / ex["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp64 (exact mode); fp32 terms (fast mode)",
        "data": "synthetic (seeded Plummer sphere)",
        "config": {"workload": w["desc"], "n": n, "theta": 0.3, "softening": 1e-4},
        "exact": ex, "fast": fa, "fp64_direct": {"ms_per_step": direct_ms},
        "roofline": {"bound": "mfma", "bound_detail": BOUND_FP64, "achieved": achieved / 1e12,
                     "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64,
                     "traffic": None, "kernel": "k_bh_force + k_bh_sun (exact tree walk)",
                     "flops_per_term": FLOPS_PER_BH_TERM, "avg_launch_ms": ex["force_ms"]},
    }))


def run_dense(a, w):
    """The collision pass on dense scenes: resolution along the dependency DAG (resolve = 0)
    against one pair at a time (resolve = 1), same pairs, bit-identical states (tests)."""
    from nbodygo_amd import NBodyGPU, gen
    scenes = []
    for n, rk in ((3000, 0.25), (20000, 0.05), (20000, 0.12), (65536, 0.02)):
        b = gen.uniform_sphere(n, seed=2, radius_k=rk)
        row = {"n": n, "radius_k": rk}
        for resolve, name in ((0, "dag"), (1, "one_lane")):
            with NBodyGPU(resolve=resolve, pair_capacity=1 << 21) as g:
                g.load(b)
                g.step(1e-3, 1.0)  # warm-up cycle
                g.load(b)
                g.set_timing(True)
                st = g.step(1e-3, 1.0)
            row["pairs"] = st["pairs_last"]
            row[name + "_resolve_ms"] = st["resolve_ms"]
            row[name + "_step_ms"] = st["wall_ms"]
        row["speedup"] = row["one_lane_resolve_ms"] / row["dag_resolve_ms"]
        log("dense n=%d: %d pairs, %.3f vs %.3f ms" % (n, row["pairs"], row["dag_resolve_ms"],
                                                       row["one_lane_resolve_ms"]))
        scenes.append(row)
    big = max(scenes, key=lambda r: r["pairs"])
    print(json.dumps({
        "metric": METRIC, "value": big["pairs"] / (big["dag_resolve_ms"] * 1e-3),
        "unit": "collision pairs resolved per second (densest scene)", "n_gpus": 1, "steps": 1,
        "warmup": 1, "ms_per_step": big["dag_step_ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp64", "data": "synthetic (seeded)",
        "config": {"workload": w["desc"]}, "scenes": scenes}))


def run_lifecycle(a, w):
    """Row (f)1: the collection bookkeeping on the device (nb_remove_bodies + nb_add_bodies +
    nb_cycle) against a host round trip (nb_store_bodies, numpy compaction, nb_load)."""
    from nbodygo_amd import NBodyGPU, gen
    from nbodygo_amd.gen import concat
    from nbodygo_amd.runner import _select
    n = w["n"]
    b = gen.plummer(n, seed=2024)
    add = gen.plummer(1000, seed=7)
    reps = max(1, a.steps // 10)
    g = NBodyGPU(G=1.0, eps=0.01, overlap_cutoff=False, collisions=False)
    dev, host = [], []
    for _ in range(reps):
        g.load(b)
        g.step(1e-3, 1.0, 1)
        t0 = time.perf_counter()
        g.remove_bodies(n // 100)
        g.add_bodies(add)
        n_dev = g.cycle(1.0)
        dev.append(time.perf_counter() - t0)
        g.load(b)
        g.step(1e-3, 1.0, 1)
        t0 = time.perf_counter()
        s = g.store()
        keep = np.ones(s.n, bool)
        keep[::100] = False
        g.load(concat(_select(s, keep), add))
        host.append(time.perf_counter() - t0)
    g.close()
    dms, hms = 1e3 * float(np.median(dev)), 1e3 * float(np.median(host))
    print(json.dumps({
        "metric": METRIC, "value": 1e3 / dms, "unit": "cycles/s (delete 1 % + 1,000 adds + Cycle)",
        "n_gpus": 1, "steps": reps, "warmup": 0, "ms_per_step": dms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp64 state", "data": "synthetic",
        "config": {"workload": w["desc"], "n": n, "bodies_after": n_dev},
        "device_ms": dms, "host_round_trip_ms": hms, "speedup": hms / dms}))


def host_cpu():
    """Model name and the number of cores this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        cores = os.cpu_count() or 1
    return model, cores


def cpu_threads():
    """Worker threads for the CPU legs: every core this process may run on, capped by the
    OMP_NUM_THREADS share the host gives it when that is set."""
    _, cores = host_cpu()
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        cores = min(cores, int(env))
    return max(1, cores)


def cpu_baseline(b, seconds):
    """The reference's worker-pool CPU algorithm (oracle restatement) on a bounded sample: the
    first T targets of the same body set against all N sources, whole Compute phases."""
    from oracle import binding as ob
    threads = cpu_threads()
    model, affinity = host_cpu()
    st = ob.State(b.x, b.y, b.z, b.vx, b.vy, b.vz, b.m, b.r)
    t = min(b.n, 256)
    el, _, _ = ob.bench_workpool(st, 1.0, t, nsteps=1, nthreads=threads)
    rate = t * (b.n - 1) / el
    t = int(min(b.n, max(threads, rate * seconds / (b.n - 1))))
    # when every target fits in the time, whole Compute phases are repeated to fill it
    phases = max(1, min(10, int(round(rate * seconds / (t * (b.n - 1))))))
    el, _, _ = ob.bench_workpool(st, 1.0, t, nsteps=phases, nthreads=threads)
    inter = phases * t * (b.n - 1)
    return {"value": inter / el / 1e9, "unit": "G interactions/s", "cores": threads,
            "kind": "port", "cpu_model": model, "affinity_cores": affinity,
            "sample": "reference worker-pool algorithm (AoS bodies via pointers, force loop + "
                      "collision-detection loop per target, %d threads over contiguous slices), "
                      "%d targets x %d sources of the headline sphere, %d Compute phase(s), "
                      "%.1f s" % (threads, t, b.n, phases, el),
            "seconds": el}


def cpu_soa(b, seconds):
    """The stronger CPU line: an optimised FP64 softened direct sum (SoA arrays, vectorised
    inner loop, OpenMP threads over target slices; not the reference's algorithm) on a bounded
    sample of the same body set: the first T targets against all N sources."""
    from oracle import binding as ob
    threads = cpu_threads()
    model, affinity = host_cpu()
    t = min(b.n, 4 * threads)
    _, el = ob.soa_soft(b.x, b.y, b.z, b.m, 1.0, 0.01, t0=0, t1=t, nthreads=threads)
    rate = t * (b.n - 1) / el
    t = int(min(b.n, max(t, rate * seconds / (b.n - 1))))
    _, el = ob.soa_soft(b.x, b.y, b.z, b.m, 1.0, 0.01, t0=0, t1=t, nthreads=threads)
    return {"value": t * (b.n - 1) / el / 1e9, "unit": "G interactions/s", "cores": threads,
            "kind": "optimised (not the reference's algorithm)", "cpu_model": model,
            "affinity_cores": affinity,
            "sample": "FP64 softened direct sum over SoA arrays, AVX2 vectorised inner loop, "
                      "%d OpenMP threads over contiguous target slices, %d targets x %d sources "
                      "of the headline sphere, %.1f s" % (threads, t, b.n, el),
            "seconds": el}


def cpu_config1():
    """The 1,000-body FP64 scene (uniform sphere of radius 1e3, lognormal masses, radii
    0.05 m^(1/3), velocities N(0, 5^2), dt 1e-3, elastic collisions, G = 6.673e-11, no
    softening) for 100 full cycles of the reference's CPU algorithm: the worker pool's Compute
    phase, then the serial resolution and Update."""
    from nbodygo_amd import gen
    from oracle import binding as ob
    threads = cpu_threads()
    b = gen.uniform_sphere(1000, seed=1)
    runs = []
    for _ in range(5):  # the same 100 cycles from the same start; the median run is reported
        st = ob.State.from_bodies(b)
        runs.append(ob.bench_cycles(st, 6.673e-11, 1e-3, 1.0, 100, threads))
    el, stats = sorted(runs, key=lambda r: r[0])[len(runs) // 2]
    n = b.n
    return {"value": 100 / el, "unit": "steps/s", "cores": threads, "kind": "port",
            "G_interactions_per_s": 100 * n * (n - 1) / el / 1e9, "seconds": el,
            "pairs_total": stats["pairs_total"], "resolved_total": stats["resolved_total"],
            "sample": "N=1000 FP64, reference semantics (eps 0, overlap cutoff), 100 cycles, "
                      "median of 5 runs: worker-pool Compute (force + detection), serial coll3d "
                      "resolution in canonical order, serial Update"}


def run_batch(a, w):
    r = batch_leg(a.steps, a.warmup, tune=not a.no_tune)
    print(json.dumps({
        "metric": METRIC, "value": r["value"], "unit": "G interactions/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": r["ms_per_step"],
        "steps_per_s": r["steps_per_s"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (seeded clusters)",
        "config": {"workload": w["desc"], "sims": 512, "bodies": r["bodies"],
                   "shapes": r["shapes"]},
        "roofline": {"bound": "mfma", "bound_detail": BOUND_FP32,
                     "achieved": r["frac_of_fp32_peak_whole_step"] * PEAK_FP32 / 1e12,
                     "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                     "frac": r["frac_of_fp32_peak_whole_step"], "traffic": None,
                     "kernel": "k_batch_force (whole-step time, update included)",
                     "flops_per_interaction": FLOPS_PER_INTERACTION}}))


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: runs on one GPU only (sharding the targets over several GPUs is not "
              "built); got --gpus %d" % a.gpus, file=sys.stderr)
        sys.exit(2)
    from nbodygo_amd import gen

    w = WORKLOADS[a.workload]
    pinned = shape_kw(a.shape)
    tune = not a.no_tune
    if a.workload == "batch512":
        return run_batch(a, w)
    if a.workload == "collide2x524k":
        return run_collide(a, w, pinned)
    if a.workload == "barneshut262k":
        return run_barneshut(a, w)
    if a.workload == "lifecycle1m":
        return run_lifecycle(a, w)
    if a.workload == "dense_collisions":
        return run_dense(a, w)
    if a.workload == "surrogate262k":
        sur = surrogate_leg(a, pinned)
        print(json.dumps({
            "metric": METRIC, "value": sur["equiv_G_interactions_per_s"],
            "unit": "G interactions/s (N(N-1) per step, surrogate)", "n_gpus": 1,
            "steps": a.sur_steps, "warmup": 3, "ms_per_step": sur["ms_per_step"],
            "steps_per_s": sur["steps_per_s"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32 (near field, expansions) + bf16 MFMA (MLP)",
            "data": "synthetic (seeded Plummer spheres)",
            "config": {"workload": w["desc"], "n": sur["n"]}, "surrogate": sur}))
        return
    n = w["n"]
    p = plain_leg(n, a.steps, a.warmup, pinned, tune)
    log("plain N=%d: %.3f ms/step, shape %s" % (n, p["ms_per_step"], p["kernel_shape"]))
    out = {
        "metric": METRIC, "value": p["value"], "unit": p["unit"], "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": p["ms_per_step"],
        "steps_per_s": p["steps_per_s"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (seeded Plummer sphere)",
        "config": {"workload": w["desc"], "n": n, "eps": 0.01, "dt": 1e-3, "G": 1.0,
                   "collisions": "off", "kernel_shape": p["kernel_shape"]},
        "roofline": p["roofline"], "force_error": p["force_error"],
    }
    if a.workload == "plummer65k" and not (a.no_elastic or a.headline_only):
        out["elastic"] = elastic_leg(n, 1e-4, a.steps, a.warmup, pinned, tune)
        out["elastic_dense"] = elastic_leg(n, 5e-3, a.steps, a.warmup, pinned, tune)
    if a.workload == "plummer262k" and not a.headline_only:
        if not a.no_surrogate:
            out["surrogate"] = surrogate_leg(a, pinned, plain=p)
            log("surrogate: %.3f ms/step" % out["surrogate"]["ms_per_step"])
        if not a.no_extras:
            c2 = plain_leg(65536, a.steps, a.warmup, pinned, tune)
            if not a.no_elastic:
                c2["elastic"] = elastic_leg(65536, 1e-4, a.steps, a.warmup, pinned, tune)
                c2["elastic_dense"] = elastic_leg(65536, 5e-3, a.steps, a.warmup, pinned, tune)
            out["plummer65k"] = c2
            log("plummer65k: %.3f ms/step" % c2["ms_per_step"])
            out["batch512"] = batch_leg(a.steps, a.warmup, tune)
            log("batch512: %.3f ms/step" % out["batch512"]["ms_per_step"])
            out["collide2x524k"] = collide_leg(a, pinned, a.c5_steps or 20)
            log("collide2x524k: %.1f x" % out["collide2x524k"]["speedup_vs_plain"])
    if not (a.no_cpu or a.headline_only):
        out["cpu_baseline"] = cpu_baseline(gen.plummer(n, seed=2024), a.cpu_seconds)
        out["cpu_soa"] = cpu_soa(gen.plummer(n, seed=2024), a.cpu_seconds / 2)
        out["cpu_config1"] = cpu_config1()
        log("cpu legs done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
