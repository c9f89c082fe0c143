"""Benchmark: OLDG node-updates/s on one MI355X, with the HBM roofline of the
stage kernel and the oracle timed on the host cores beside it.

Workload (default): the 3D periodic uniform forest, P = 4, E = 4, roots = 16
(4096 blocks of 20^3 nodes, 320^3 = 32.8M nodes per field), standing mode
u = sin(2 pi x) sin(2 pi y) sin(2 pi z), v = 0, classic RK4 with the truncation
as stage and step limiter, dt = 0.5 h / (P+1).  A node-update is one node of
one field advanced by one full RK4 step (4 right-hand sides, 4 truncations and
the stage arithmetic), the reference's "ns per node and step".

Timed region: K RK4 steps with the state resident in HBM, bracketed by device
synchronizations.  Printed: one JSON line on stdout; progress on stderr.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision f64|f32c]
                  [--accuracy-steps S] [--no-cpu]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "treewaveoldg.jl_amd"))

PEAK_HBM_GBS = 8000.0
# algorithmic values moved per node-update by the fused RK4 stages.  The
# compensated form keeps three registers: stage 1 reads y_n (2), writes the
# accumulator's v (1) and y_2 (2); stage 2 reads y_2, y_n (4) and the
# accumulator's v (1), writes acc, y_3 (4) (its accumulator u, v_n + 2 k2_u, from
# y_n's v); stage 3 reads y_3, y_n, acc (6) and writes acc, y_4 (4); stage 4
# reads y_4, y_n, acc and the low words of u and v (8) and writes y_{n+1} and the
# low words (4): 36 values.  FP64 keeps no accumulator: stage 1 reads y_n (2) and
# writes y_2 (2); stage 2 reads y_2, y_n (4) and writes y_3 (2); stage 3 reads
# y_3, y_2, y_n (6) and writes y_4 and z = (y_2 + 2 y_3)/3 + dt/3 k3 (4); stage 4
# reads y_4 and z (4) and writes y_{n+1} (2): 26 values.
VALUES_PER_NODE_UPDATE = {"f64": 26, "f32c": 36}
BYTES_PER_VALUE = {"f64": 8, "f32c": 4}
# The split form (u in FP64, v in FP32) moves u's 13 values in FP64, without an
# accumulator as in FP64, and v's 17 in FP32 with the accumulator, plus the low
# word of v read and written at stage 4.
BYTES_PER_NODE_UPDATE = {"f64": 208, "f32c": 144, "split": 13 * 8 + 17 * 4 + 2 * 4}


# HBM bytes per launch of the stage kernel on the default workload, measured with
# rocprofv3 (FETCH_SIZE and WRITE_SIZE, scaled by a device copy's reading of the
# same counters); the summary holding them sits beside this file.
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_fused3_periodic320_p4e4.txt")


def measured_traffic(precision):
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                w = line.split()
                if len(w) == 3 and w[0] == "traffic_bytes_per_launch" and w[1] == precision:
                    return int(w[2])
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_problem(precision, D=3, P=4, E=4, roots=16):
    import twoldg as T
    forest = T.wave_forest(D, P, E, roots=roots)
    prob = T.WaveProblem(forest, T.Element(P), precision=precision)
    return T, forest, prob


def gpu_run(args, precision):
    T, forest, prob = build_problem(precision)
    nodes = forest.nblocks * forest.N ** 3
    log(f"[bench] {precision}: {forest.nblocks} blocks, {nodes} nodes per field, "
        f"device bytes {prob.info()['device_bytes'] / 2**30:.2f} GiB")
    u0 = T.initial_state(prob, T.wave_exact(3, 1.0, 1, 0.0))
    dt = T.oldg_timestep(prob, 0.5)
    total = args.warmup + args.steps
    integ = T.wave_integrator(prob, u0, 0.0, total * dt, dt)
    integ.step(args.warmup)
    T._lib.check(T._lib.lib().twoldg_synchronize(prob._h))
    t0 = time.perf_counter()
    ms, kernel_ms, launches = integ.timed(args.steps)
    T._lib.check(T._lib.lib().twoldg_synchronize(prob._h))
    wall = time.perf_counter() - t0
    value = nodes * args.steps / wall
    log(f"[bench] {precision}: {args.steps} steps in {wall * 1e3:.1f} ms (events {ms:.1f} ms), "
        f"stage kernels {kernel_ms:.1f} ms over {launches} launches -> {value:.3e} node-updates/s")
    alg_bytes = BYTES_PER_NODE_UPDATE[precision] * nodes * args.steps
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9 if kernel_ms > 0 else None
    res = dict(value=value, ms_per_step=wall * 1e3 / args.steps, kernel_ms=kernel_ms,
               launches=launches, achieved_gbs=achieved,
               bytes_per_launch=alg_bytes / max(launches, 1), nodes=nodes, dt=dt)
    if args.accuracy_steps:
        res.update(accuracy(T, prob, args.accuracy_steps, dt))
    prob.close()
    return res


def accuracy(T, prob, nsteps, dt):
    """L2 error against the exact standing mode after nsteps steps (the 320^3 workload's
    200), and the discrete energy drift."""
    u0 = T.initial_state(prob, T.wave_exact(3, 1.0, 1, 0.0))
    e0 = T.wave_energy(prob, u0)
    integ = T.wave_integrator(prob, u0, 0.0, nsteps * dt, dt)
    integ.solve()
    e1 = integ.energy()
    exact = T.fill_at_nodes(T.wave_exact(3, 1.0, 1, nsteps * dt), prob)
    prob.upload(exact, T._lib.SLOT_AUX)
    from twoldg.drivers import _errors
    l2, linf = _errors(prob, (1, 2))
    log(f"[bench] {prob.precision}: {nsteps} steps, l2 error {l2:.4e}, linf {linf:.4e}, "
        f"energy drift {(e1 - e0) / e0:.3e}")
    return dict(l2=l2, linf=linf, energy_drift=(e1 - e0) / e0, accuracy_steps=nsteps)


def cpu_baseline(seconds=15.0):
    """The oracle (the CPU FP64 restatement of the same algorithm) on this host's
    cores: RK4 steps of the same workload until ~`seconds` of CPU work."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = O.lib().orc_threads()
    pr = O.Problem.brick(3, 4, 4, roots=16)
    y = pr.truncate(pr.fill_exact(0, 0.0, (1.0, 1)))
    dt = pr.timestep(0.5)
    nodes = pr.nb * pr.nn
    t0 = time.perf_counter()
    y = pr.rk4(y, 0.0, dt, 1)  # first step: first touch of every work array
    first = time.perf_counter() - t0
    steps = max(1, min(20, int(seconds / max(first, 1e-3))))
    t0 = time.perf_counter()
    y = pr.rk4(y, dt, dt, steps)
    el = time.perf_counter() - t0
    value = nodes * steps / el
    log(f"[bench] cpu oracle: {threads} threads, {steps} steps in {el:.2f} s "
        f"-> {value:.3e} node-updates/s ({el / steps / nodes * 1e9:.2f} ns per node and step)")
    return dict(value=value, unit="node-updates/s", cores=threads, kind="port",
                sample=f"{steps} RK4 steps of the same 32.8M-node forest after one warm-up "
                       f"step, FP64 C++/OpenMP restatement (oracle), {threads} threads",
                host=host_info(), ns_per_node_update=el / steps / nodes * 1e9)


def host_info():
    """The host the CPU baseline ran on: CPU model, logical CPUs visible to this
    process, NUMA nodes."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    numa = None
    try:
        numa = len([d for d in os.listdir("/sys/devices/system/node") if d.startswith("node")])
    except OSError:
        pass
    return dict(cpu_model=model, nproc=len(os.sched_getaffinity(0)), numa_nodes=numa)


def timed_integrator(T, prob, u0, dt, steps, warmup=2):
    """Wall time of `steps` RK4 steps in one call (the library replays its
    captured step), after `warmup` steps in one call (which captures it)."""
    integ = T.wave_integrator(prob, u0, 0.0, (steps + warmup) * dt, dt)
    integ.step(warmup)
    T._lib.check(T._lib.lib().twoldg_synchronize(prob._h))
    t0 = time.perf_counter()
    integ.step(steps)
    T._lib.check(T._lib.lib().twoldg_synchronize(prob._h))
    wall = time.perf_counter() - t0
    return wall, None, None, integ


def periodic2d():
    """wave_errors(Val(2); P=4, E=2, roots=8) end to end on the GPU (parity run)."""
    import twoldg as T
    t0 = time.perf_counter()
    r = T.wave_errors(2, 4, 2, roots=8)
    el = time.perf_counter() - t0
    nodes = r["nblocks"] * 100
    # the same 142 steps again, timed alone (graph replays of the batch stage)
    prob = r["problem"]
    u0 = T.initial_state(prob, T.wave_exact(2, 1.0, 1, 0.0))
    wall, _, _, _ = timed_integrator(T, prob, u0, r["dt"], r["nsteps"])
    log(f"[bench] 2D periodic 80^2: l2 {r['l2']:.4e} (u {r['l2u']:.3e}, v {r['l2v']:.3e}), "
        f"E1/E0 {r['energy'] / r['energy0']:.12f}, {r['nsteps']} steps, {el:.2f} s wall with "
        f"setup; steps alone {wall * 1e6 / r['nsteps']:.1f} us per step "
        f"({nodes * r['nsteps'] / wall:.3e} node-updates/s), stages {prob.info()['stages']}")
    return dict(l2=r["l2"], l2u=r["l2u"], l2v=r["l2v"], linf=r["linf"],
                energy_ratio=r["energy"] / r["energy0"], nsteps=r["nsteps"], wall_s=el,
                us_per_step=wall * 1e6 / r["nsteps"], value=nodes * r["nsteps"] / wall,
                nodes=nodes, note="6400 nodes: launch-bound, a parity run")


def nested_cubes(large=False, steps=500):
    """Reflecting box with nested cubes (3 levels, finest half-width 1/8): P=3,
    E=2, roots=8 (1408 blocks), 500 steps at the finest level's CFL (0.5); or
    the large refined forest P=4, E=4, roots=16 (11264 blocks of 20^3) for the
    bandwidth figure.  Cos mode m=2, FP64 and FP32c; fused stages with the
    coarse-fine face ghosts formed by two face-sized launches per stage.
    Reported: throughput over the steps after a two-step warm-up, the energy
    drift over all steps, and FP32c's L2 distance from the FP64 run."""
    import twoldg as T
    P, E, roots = (4, 4, 16) if large else (3, 2, 8)
    out = {}
    forest = T.wave_forest(3, P, E, roots=roots, reflecting=True,
                           refined=T.nested_cubes(3, 1 / 8, 3))
    nodes = forest.nblocks * forest.N ** 3
    warm = 2
    final = {}
    for prec in ("f64", "f32c"):
        prob = T.WaveProblem(forest, T.Element(P), precision=prec)
        u0 = T.initial_state(prob, T.reflect_exact(3, 1.0, 2, 0.0))
        dt = T.oldg_timestep(prob, 0.5)
        e0 = T.wave_energy(prob, u0)
        wall, kms, nl, integ = timed_integrator(T, prob, u0, dt, steps - warm, warmup=warm)
        e1 = integ.energy()
        timed = steps - warm
        alg = BYTES_PER_NODE_UPDATE[prec] * nodes * timed
        final[prec] = integ.u
        # the error against the exact cos standing mode at the end (reflect_exact)
        exact = T.fill_at_nodes(T.reflect_exact(3, 1.0, 2, steps * dt), prob)
        prob.upload(exact, T._lib.SLOT_AUX)
        from twoldg.drivers import _errors
        l2, linf = _errors(prob, (1, 2))
        out[prec] = dict(value=nodes * timed / wall, ms_per_step=wall * 1e3 / timed,
                         alg_gbs_whole_step=alg / wall / 1e9, energy_drift=(e1 - e0) / e0,
                         steps=steps, l2_vs_exact=l2, linf_vs_exact=linf)
        log(f"[bench] nested cubes{' (large)' if large else ''} {prec}: {forest.nblocks} blocks, "
            f"{steps} steps, {nodes * timed / wall:.3e} node-updates/s, "
            f"{alg / wall / 1e9:.0f} GB/s algorithmic, energy drift {(e1 - e0) / e0:.2e}, "
            f"L2 vs exact {l2:.4e}")
        prob.close()
    rel = float(np.linalg.norm(final["f32c"] - final["f64"]) / np.linalg.norm(final["f64"]))
    out["f32c_vs_f64_rel_l2"] = rel
    out["f32c_over_f64_l2_vs_exact"] = out["f32c"]["l2_vs_exact"] / out["f64"]["l2_vs_exact"]
    log(f"[bench] nested cubes{' (large)' if large else ''}: FP32c vs FP64 after {steps} steps, "
        f"relative L2 {rel:.2e}; L2 vs exact FP32c / FP64 "
        f"{out['f32c_over_f64_l2_vs_exact']:.3f}")
    out["nodes"] = nodes
    out["blocks"] = forest.nblocks
    return out


def batch2d(steps=20):
    """512 independent 2D periodic problems, P in {2..6}, roots in {4, 8, 16},
    E=2, seeded Gaussian pulses, one launch per stage, graph-replayed steps."""
    import twoldg as T
    from twoldg.batch import gaussian_pulses, initial_states
    probs = gaussian_pulses()
    u0s = initial_states(probs)
    out = {}
    final = {}
    for prec in ("f32c", "f64"):
        t0 = time.perf_counter()
        batch = T.Batch([p[:3] for p in probs], precision=prec)
        setup = time.perf_counter() - t0
        for i, u in enumerate(u0s):
            batch.upload(i, u)
        batch.rk4(2)
        ms = batch.rk4(steps)
        info = batch.info()
        final[prec] = [batch.download(i) for i in range(len(probs))]
        nodes = info["nodes"]
        alg = BYTES_PER_NODE_UPDATE[prec] * nodes * steps
        out[prec] = dict(value=nodes * steps / (ms * 1e-3), ms_per_step=ms / steps,
                         alg_gbs=alg / (ms * 1e-3) / 1e9, tiles=info["tiles"],
                         workgroups_per_launch=info["workgroups"], setup_s=setup)
        log(f"[bench] 2D batch {prec}: {len(probs)} problems, {nodes} nodes, "
            f"{nodes * steps / (ms * 1e-3):.3e} node-updates/s ({ms / steps:.3f} ms/step, "
            f"{alg / (ms * 1e-3) / 1e9:.0f} GB/s), tiles {info['tiles']}")
        batch.close()
    # FP32c against FP64 after the same 2 + steps steps: over the whole batch and
    # the worst problem (relative L2)
    a32 = np.concatenate(final["f32c"])
    a64 = np.concatenate(final["f64"])
    out["f32c_vs_f64_rel_l2"] = float(np.linalg.norm(a32 - a64) / np.linalg.norm(a64))
    out["f32c_vs_f64_rel_l2_worst_problem"] = max(
        float(np.linalg.norm(x - y) / np.linalg.norm(y)) for x, y in zip(final["f32c"], final["f64"]))
    log(f"[bench] 2D batch: FP32c vs FP64 after {steps + 2} steps, relative L2 "
        f"{out['f32c_vs_f64_rel_l2']:.2e} (worst problem "
        f"{out['f32c_vs_f64_rel_l2_worst_problem']:.2e})")
    out["nodes"] = nodes
    out["problems"] = len(probs)
    out["seed"] = 20261203
    return out


def ball(steps=20):
    """Seven-patch ball, P=3, E=2, roots=(4,4) (448 blocks), plane wave
    k = 2 pi (1,1,1)/sqrt(3), phi = 0, Dirichlet sphere; FP64 and FP32c; l2 at
    t = 1 for both forms."""
    import twoldg as T
    k = tuple(2 * math.pi / math.sqrt(3) for _ in range(3))
    out = {}
    for prec in ("f64", "f32c"):
        r = T.sevenpatch_errors(3, 2, roots=(4, 4), k=k, phi=0.0, precision=prec)
        prob = r["problem"]
        nodes = prob.nblocks * prob.nn
        u0 = T.initial_state(prob, T.sevenpatch_exact(T.PlaneWave(k, 0.0), 0.0))
        wall, kms, nl, integ = timed_integrator(T, prob, u0, r["dt"], steps)
        out[prec] = dict(l2=r["l2"], l2u=r["l2u"], l2v=r["l2v"], linf=r["linf"],
                         energy_ratio=r["energy"] / r["energy0"], nsteps=r["nsteps"],
                         value=nodes * steps / wall, ms_per_step=wall * 1e3 / steps)
        log(f"[bench] seven-patch ball {prec}: {prob.nblocks} blocks, l2 {r['l2']:.4e} at t=1 "
            f"({r['nsteps']} steps), E1/E0 {r['energy'] / r['energy0']:.6f}, "
            f"{nodes * steps / wall:.3e} node-updates/s")
        prob.close()
    out["f32c_over_f64_l2"] = out["f32c"]["l2"] / out["f64"]["l2"]
    out["glue_energy_balance"] = glue_energy_balance(T)
    return out


def glue_energy_balance(T, P=3, E=2, roots=(4, 4), seed=5):
    """The energy flux across the seven-patch ball's glues (sevenpatch_tests.jl:
    182-226) on the benchmarked ball, FP64 and FP32c: zero Dirichlet data and a
    seeded random state that is zero in the outer shells' outermost elements, so
    that nothing crosses the sphere; the rate of the discrete energy along the
    right-hand side, (E(y + e f) - E(y - e f)) / 2e (exact for the quadratic
    energy), is then the glues' net flux, reported against the Cauchy-Schwarz
    bound 2 sqrt(E(y) E(f)) of its terms."""
    forest, geo = T.ball_forest(P, E, roots=roots)
    N = forest.N
    faces = forest.faces()
    res = {}
    for prec in ("f64", "f32c"):
        prob = T.WaveProblem(forest, T.Element(P), precision=prec, geometry=geo,
                             boundary=T.ZeroData())
        y = np.random.default_rng(seed).standard_normal(prob.state_len).reshape(
            forest.nblocks, 2, N, N, N)
        for b in range(forest.nblocks):
            for f in range(6):
                if faces[b * 6 + f].kind != T._lib.FACE_DIRICHLET:
                    continue
                d, side = divmod(f, 2)
                idx = [slice(None)] * 3
                idx[2 - d] = slice(N - P - 1, N) if side else slice(0, P + 1)
                y[b][(slice(None),) + tuple(idx)] = 0.0
        y = T.truncate_state(y.reshape(-1), prob)
        f = np.empty_like(y)
        T.wave_rhs(f, y, prob, 0.0)
        ey, ef = T.wave_energy(prob, y, t=0.0), T.wave_energy(prob, f, t=0.0)
        eps = math.sqrt(ey / ef)
        rate = (T.wave_energy(prob, y + eps * f, t=0.0) -
                T.wave_energy(prob, y - eps * f, t=0.0)) / (2 * eps)
        bound = 2 * math.sqrt(ey * ef)
        res[prec] = dict(rate=rate, bound=bound, rel=abs(rate) / bound)
        log(f"[bench] seven-patch ball glues {prec}: energy rate {rate:.3e} against the "
            f"bound {bound:.3e} of its terms ({abs(rate) / bound:.1e})")
        prob.close()
    return res


def annulus(s=8):
    """The four-patch annulus (diagonal metric, identity glues, walls), P=3, E=2,
    roots (2s, s) per patch (512 blocks at s=8), the rotating mode m=2, n=1 to a
    quarter turn; FP64 and FP32c, l2 error and throughput over the run."""
    import twoldg as T
    out = {}
    for prec in ("f64", "f32c"):
        t0 = time.perf_counter()
        r = T.annulus_errors(3, 2, roots=(2 * s, s), m=2, precision=prec)
        el = time.perf_counter() - t0
        prob = r["problem"]
        nodes = prob.nblocks * prob.nn
        u0 = T.initial_state(prob, T.annulus_exact(prob.geometry, prob.forest, 2, r["k"], 0.0))
        wall, _, _, _ = timed_integrator(T, prob, u0, r["dt"], 50)
        out[prec] = dict(l2=r["l2"], l2u=r["l2u"], l2v=r["l2v"], nsteps=r["nsteps"],
                         energy_loss=1 - r["energy"] / r["energy0"],
                         value=nodes * 50 / wall, driver_wall_s=el)
        log(f"[bench] annulus s={s} {prec}: {prob.nblocks} blocks, l2 u {r['l2u']:.3e} "
            f"v {r['l2v']:.3e} ({r['nsteps']} steps), 1 - E1/E0 {1 - r['energy'] / r['energy0']:.2e}, "
            f"{nodes * 50 / wall:.3e} node-updates/s")
        prob.close()
    return out


# The reference's own GPU benchmark figures (CODE.md:5108-5122, one H200,
# ns per node and RK4 step, the least of five): its flat box and its ball.
REFERENCE_H200_NS = {"flatbox16": {"f64": 1.61, "f32c": 1.26},
                     "benchmark-ball": {"f64": 2.84, "f32c": 1.92}}


def flatbox16(steps=50):
    """The reference's flat-box benchmark (bin/benchmark.jl, CODE.md:4998-5007):
    3D periodic, P = 3, 4096 blocks of 16^3 (E = 4, roots = 16), 16.8M nodes,
    standing sine mode, truncation on; ns per node and step beside the
    reference's published H200 figure."""
    import twoldg as T
    forest = T.wave_forest(3, 3, 4, roots=16)
    nodes = forest.nblocks * forest.N ** 3
    out = {}
    for prec in ("f64", "f32c"):
        prob = T.WaveProblem(forest, T.Element(3), precision=prec)
        u0 = T.initial_state(prob, T.wave_exact(3, 1.0, 1, 0.0))
        dt = T.oldg_timestep(prob, 0.5)
        wall, _, _, _ = timed_integrator(T, prob, u0, dt, steps)
        ns = wall / steps / nodes * 1e9
        ref = REFERENCE_H200_NS["flatbox16"][prec]
        out[prec] = dict(value=nodes * steps / wall, ns_per_node_step=ns,
                         reference_h200_ns=ref, speedup_vs_reference_h200=ref / ns)
        log(f"[bench] flat box 4096 x 16^3 P=3 {prec}: {ns:.4f} ns per node and step "
            f"(reference H200 {ref}), {nodes * steps / wall:.3e} node-updates/s")
        prob.close()
    out["nodes"] = nodes
    return out


def benchmark_ball(steps=10):
    """The reference's benchmark ball (CODE.md:5056-5066): the thirteen-patch
    ball Lc = 4, Rm = 12, R1 = 100, roots (4, 4, 44), P = 3, E = 4 (4672 blocks
    of 16^3, 1.91e7 nodes), radiative sphere with u/R, truncation on (per-element
    tables on the inner shells, factored on the outer ones); an outgoing pulse of
    width 1; ns per node and step beside the reference's published H200 figure."""
    import twoldg as T
    t0 = time.perf_counter()
    forest, geo = T.thirteenpatch_forest(3, 4, roots=(4, 4, 44), Lc=4.0, Rm=12.0, R1=100.0)
    nodes = forest.nblocks * forest.N ** 3
    out = {}
    for prec in ("f64", "f32c"):
        prob = T.WaveProblem(forest, T.Element(3), geometry=geo, precision=prec,
                             boundary=T.RadiativeBoundary())
        build = time.perf_counter() - t0
        from twoldg.thirteenpatch import pulse_state
        u0 = T.initial_state(prob, pulse_state(T.OutgoingPulse(1.0), 0.0))
        dt = T.oldg_timestep(prob, 0.5)
        wall, _, _, _ = timed_integrator(T, prob, u0, dt, steps)
        ns = wall / steps / nodes * 1e9
        ref = REFERENCE_H200_NS["benchmark-ball"][prec]
        out[prec] = dict(value=nodes * steps / wall, ns_per_node_step=ns, dt=dt,
                         ms_per_step=wall * 1e3 / steps, reference_h200_ns=ref,
                         speedup_vs_reference_h200=ref / ns, build_s=build)
        log(f"[bench] benchmark ball (13 patches, 4672 x 16^3) {prec}: {ns:.4f} ns per node and "
            f"step (reference H200 {ref}), {wall * 1e3 / steps:.2f} ms per step, dt {dt:.3e}, "
            f"host build {build:.1f} s")
        prob.close()
        t0 = time.perf_counter()
    out["nodes"] = nodes
    out["blocks"] = forest.nblocks
    return out


# The side workloads by name, each at the size the default run reports.
WORKLOADS = {
    "periodic2d": lambda: periodic2d(),
    "nested-cubes": lambda: nested_cubes(),
    "nested-cubes-large": lambda: nested_cubes(large=True, steps=10),
    "batch2d": lambda: batch2d(),
    "ball": lambda: ball(),
    "annulus": lambda: annulus(),
    "flatbox16": lambda: flatbox16(),
    "benchmark-ball": lambda: benchmark_ball(),
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="f64", choices=["f64", "f32c", "split"])
    ap.add_argument("--accuracy-steps", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--single", action="store_true", help="time only --precision's form")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the side workloads (reported beside the main line)")
    ap.add_argument("--workload", choices=sorted(WORKLOADS),
                    help="run only this side workload and print its result (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for "
              f"--gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)
    if args.workload:
        print(json.dumps(WORKLOADS[args.workload](), default=float))
        return
    main_res = gpu_run(args, args.precision)
    others = {} if args.single else {
        f: gpu_run(args, f) for f in ("f64", "f32c", "split") if f != args.precision}
    configs = {}
    if not args.no_configs:
        for name, run in WORKLOADS.items():
            configs[name] = run()
    cpu = None if args.no_cpu else cpu_baseline()
    prec = args.precision
    launch_ms = main_res["kernel_ms"] / max(main_res["launches"], 1)
    roofline = None
    if main_res["achieved_gbs"]:
        roofline = dict(bound="hbm", achieved=round(main_res["achieved_gbs"], 1),
                        peak=PEAK_HBM_GBS, unit="GB/s",
                        frac=round(main_res["achieved_gbs"] / PEAK_HBM_GBS, 4),
                        traffic=measured_traffic(prec),
                        kernel="fused RK4 stage (k_fused3)",
                        avg_launch_ms=round(launch_ms, 4),
                        alg_bytes_per_launch=round(main_res["bytes_per_launch"]))
    out = {
        "metric": "OLDG node-updates/s + % of HBM BW, FP32 batched vs FP64 at equal L2 error",
        "value": main_res["value"],
        "unit": "node-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": prec,
        "data": "synthetic",
        "config": {"workload": "3D periodic uniform forest, P=4, E=4, roots=16 "
                               "(4096 blocks of 20^3, 32.8M nodes per field), standing sine "
                               "mode, RK4 with truncation, dt = 0.5 h/(P+1)",
                   "global_batch": main_res["nodes"], "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "forms": {f: {k: r.get(k) for k in ("value", "l2", "linf", "energy_drift", "achieved_gbs")}
                  for f, r in [(prec, main_res)] + list(others.items())},
    }
    if configs:
        out["configs"] = configs
    allres = dict(others)
    allres[prec] = main_res
    if "f64" in allres and allres["f64"].get("l2"):
        f64 = allres["f64"]
        for f in ("f32c", "split"):
            if f in allres and allres[f].get("l2"):
                out["forms"][f + "_over_f64_l2"] = allres[f]["l2"] / f64["l2"]
                out["forms"][f + "_over_f64_speed"] = allres[f]["value"] / f64["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
