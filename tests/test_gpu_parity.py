"""GPU parity: the HIP library against the oracle on the same inputs.

Tolerances follow the reference's device tests (device_tests.jl:70-74):
rtol = sqrt(eps) on l2 and energy, ||delta|| <= sqrt(eps) ||host|| on arrays.
A single RHS or truncation in FP64 is held much tighter (1e-12 of the array's
scale): it differs from the oracle only by summation order.
"""
import math

import numpy as np
import pytest

import oracle as O
import twoldg as T

pytestmark = pytest.mark.gpu

SQRT_EPS = math.sqrt(np.finfo(np.float64).eps)


def _pair(D, P, E, roots, reflecting=False, regions=(), truncate=True, precision="f64"):
    forest = T.wave_forest(D, P, E, roots=roots, reflecting=reflecting,
                           refined=list(regions) if regions else False)
    prob = T.WaveProblem(forest, T.Element(P), truncate=truncate, precision=precision)
    ref = O.Problem.brick(D, P, E, roots=roots, reflecting=reflecting, truncate=truncate,
                          regions=regions)
    return prob, ref


CASES = [
    dict(D=1, P=3, E=2, roots=4),
    dict(D=2, P=4, E=2, roots=8),                       # the 80^2 periodic forest of the published table
    dict(D=2, P=2, E=2, roots=4, reflecting=True),
    dict(D=3, P=3, E=2, roots=2),
    dict(D=3, P=4, E=4, roots=2),                       # the 20^3 blocks of the 320^3 workload
    dict(D=2, P=3, E=2, roots=4, regions=[((0.25, 0.75),) * 2]),
    dict(D=3, P=3, E=2, roots=8, reflecting=True, regions=O.nested_cubes(3, 1 / 8, 3)),
]


def _id(c):
    return "D{D}P{P}E{E}r{roots}".format(**c) + ("refl" if c.get("reflecting") else "") + \
        ("ref%d" % len(c["regions"]) if c.get("regions") else "")


@pytest.mark.parametrize("case", CASES, ids=_id)
def test_rhs_matches_oracle(case, rough):
    prob, ref = _pair(**case)
    u = rough(prob.state_len, seed=3)
    du = np.empty_like(u)
    T.wave_rhs(du, u, prob, 0.0)
    want = ref.rhs(u)
    scale = np.max(np.abs(want))
    assert np.max(np.abs(du - want)) <= 1e-12 * scale


@pytest.mark.parametrize("case", CASES[:5], ids=_id)
def test_truncate_matches_oracle(case, rough):
    prob, ref = _pair(**case)
    u = rough(prob.state_len, seed=4)
    got = T.truncate_state(u.copy(), prob)
    want = ref.truncate(u)
    assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))


@pytest.mark.parametrize("case", CASES, ids=_id)
def test_energy_and_norms_match_oracle(case, rough):
    prob, ref = _pair(**case)
    u = rough(prob.state_len, seed=5)
    e = T.wave_energy(prob, u)
    assert e == pytest.approx(ref.energy(u), rel=1e-13)
    l2 = T.quadrature_norm(prob, u)
    li = T.quadrature_norm(prob, u, math.inf)
    l2r, lir = ref.norms(u)
    assert l2 == pytest.approx(l2r, rel=1e-14)
    assert li == lir
    l2v = T.quadrature_norm(prob, u, 2, vars=(2,))
    assert l2v == pytest.approx(ref.norms(u, vars=(2,))[0], rel=1e-14)


def test_rhs_is_pure(rough):
    """The RHS never writes u and keeps no memory (evolution_tests.jl:61-75)."""
    prob, _ = _pair(2, 3, 2, 4)
    u = rough(prob.state_len, seed=6)
    u_copy = u.copy()
    a = np.empty_like(u)
    b = np.empty_like(u)
    T.wave_rhs(a, u, prob, 0.0)
    T.wave_rhs(np.empty_like(u), rough(prob.state_len, seed=7), prob, 0.0)
    T.wave_rhs(b, u, prob, 0.0)
    assert np.array_equal(u, u_copy)
    assert np.array_equal(a, b)


def test_wave_errors_2d_p4_matches_oracle():
    """wave_errors(Val(2); P=4, E=2, roots=8): the published 4.54e-5 (u 6.61e-6,
    v 6.38e-5, CODE.md:4142) and the oracle's run to sqrt(eps)."""
    r = T.wave_errors(2, 4, 2, roots=8)
    o = O.wave_errors(2, 4, 2, roots=8)
    assert r["nsteps"] == 142
    assert r["l2"] == pytest.approx(o["l2"], rel=SQRT_EPS)
    assert r["energy"] == pytest.approx(o["energy"], rel=SQRT_EPS)
    assert r["l2u"] == pytest.approx(6.61e-6, rel=5e-3)
    assert r["l2v"] == pytest.approx(6.38e-5, rel=5e-3)
    assert r["energy"] <= r["energy0"]
    u = r["integrator"].u
    assert np.linalg.norm(u - o["y"]) <= SQRT_EPS * np.linalg.norm(o["y"])


@pytest.mark.parametrize("case", [
    dict(D=2, P=2, E=2, roots=4, reflecting=True),
    dict(D=2, P=3, E=2, roots=4, regions=[((0.25, 0.75),) * 2]),
    dict(D=3, P=3, E=2, roots=8, reflecting=True, regions=O.nested_cubes(3, 1 / 8, 3)),
], ids=_id)
def test_rk4_matches_oracle(case):
    prob, ref = _pair(**case)
    # a 2D uniform problem steps by the batch stage, a refined one by the generic
    # stages, the 3D refined forest by the ghosted fused stage
    want = {2: "generic" if case.get("regions") else "batch2d", 3: "fused3d"}[case["D"]]
    assert prob.info()["stages"] == want
    refl = case.get("reflecting", False)
    f = T.reflect_exact if refl else T.wave_exact
    u0 = T.initial_state(prob, f(case["D"], 1.0, 2 if refl else 1, 0.0))
    y0 = ref.truncate(ref.fill_exact(1 if refl else 0, 0.0, (1.0, 2 if refl else 1)))
    assert np.max(np.abs(u0 - y0)) <= 1e-14
    dt = T.oldg_timestep(prob, 0.5)
    assert dt == ref.timestep(0.5)
    n = 10
    integ = T.wave_integrator(prob, u0, 0.0, n * dt, dt)
    integ.solve()
    y = ref.rk4(y0, 0.0, dt, n)
    assert np.linalg.norm(integ.u - y) <= SQRT_EPS * np.linalg.norm(y)
    assert integ.energy() == pytest.approx(ref.energy(y), rel=SQRT_EPS)


def test_pure_entry_points_leave_the_integrator_alone(rough):
    """wave_rhs, truncate_state, wave_energy and the norms between an
    integrator's steps leave its state as it was: k steps, the pure calls on
    other vectors, k more steps equal 2k uninterrupted steps bitwise, on the
    fused and the generic path, FP64 and FP32c (the integrator owns its u,
    stepping.jl:72-78; wave_rhs! touches only p's scratch, evolution.jl:1367-1371)."""
    for D, P, E, roots in [(2, 3, 2, 4), (3, 3, 2, 2)]:
        for prec in ("f64", "f32c"):
            forest = T.wave_forest(D, P, E, roots=roots)
            a = T.WaveProblem(forest, T.Element(P), precision=prec)
            b = T.WaveProblem(forest, T.Element(P), precision=prec)
            u0 = T.initial_state(a, T.wave_exact(D, 1.0, 1, 0.0))
            dt = T.oldg_timestep(a, 0.5)
            ia = T.wave_integrator(a, u0, 0.0, 8 * dt, dt)
            ib = T.wave_integrator(b, u0, 0.0, 8 * dt, dt)
            ia.step(4)
            other = rough(a.state_len, seed=21)
            du = np.empty_like(other)
            T.wave_rhs(du, other, a, 0.1)
            T.truncate_state(other.copy(), a)
            T.wave_energy(a, other)
            T.quadrature_norm(a, other)
            ia.step(4)
            ib.step(8)
            assert np.array_equal(ia.u, ib.u), (D, prec)
            # a second integrator on a takes the state slot over; the first refuses
            ic = T.wave_integrator(a, u0, 0.0, 8 * dt, dt)
            with pytest.raises(RuntimeError, match="another integrator"):
                ia.step(1)
            ic.step(1)


def test_f32c_copy_out_of_state_is_the_compensated_value():
    """FP32c: a device copy out of the state slot holds hi - lo rounded to FP32,
    the download of the state slot rounded; a copy into the state slot starts
    with no compensation."""
    import ctypes
    forest = T.wave_forest(2, 4, 2, roots=4)
    p = T.WaveProblem(forest, T.Element(4), precision="f32c")
    u0 = T.initial_state(p, T.wave_exact(2, 1.0, 1, 0.0))
    dt = T.oldg_timestep(p, 0.5)
    integ = T.wave_integrator(p, u0, 0.0, 10 * dt, dt)
    integ.step(10)
    state = integ.u
    L = T._lib.lib()
    T._lib.check(L.twoldg_state_copy(p._h, T._lib.SLOT_AUX, T._lib.SLOT_STATE))
    copy = p.download(T._lib.SLOT_AUX)
    assert np.array_equal(copy, state.astype(np.float32).astype(np.float64))
    assert not np.array_equal(copy, state)  # the state carries a compensation
    T._lib.check(L.twoldg_state_copy(p._h, T._lib.SLOT_STATE, T._lib.SLOT_AUX))
    assert np.array_equal(p.download(T._lib.SLOT_STATE), copy)


FUSED = [
    dict(D=3, P=3, E=2, roots=2),
    dict(D=3, P=2, E=2, roots=4, reflecting=True),
    dict(D=3, P=4, E=4, roots=2),                       # the 320^3 workload's 20^3 blocks
    dict(D=3, P=4, E=2, roots=3),
    dict(D=3, P=1, E=4, roots=2, reflecting=True),
    dict(D=3, P=8, E=2, roots=2),                       # the largest order, 18^3 blocks
    dict(D=3, P=5, E=4, roots=2, reflecting=True),      # 24^3 blocks, 576 threads
    dict(D=3, P=3, E=4, roots=2),                       # the flat box benchmark's 16^3 blocks
]


@pytest.mark.parametrize("case", FUSED, ids=_id)
def test_fused_stages_match_oracle(case):
    """The one-launch fused stages (faces all same-level or mirror) against the
    oracle's RK4, and against the generic exchange path."""
    forest = T.wave_forest(case["D"], case["P"], case["E"], roots=case["roots"],
                           reflecting=case.get("reflecting", False))
    fused = T.WaveProblem(forest, T.Element(case["P"]))
    generic = T.WaveProblem(forest, T.Element(case["P"]), generic=True)
    assert fused.info()["fused"] and not generic.info()["fused"]
    ref = O.Problem.brick(case["D"], case["P"], case["E"], roots=case["roots"],
                          reflecting=case.get("reflecting", False))
    refl = case.get("reflecting", False)
    y0 = ref.truncate(ref.fill_exact(1 if refl else 0, 0.0, (1.0, 2 if refl else 1)))
    # a rough perturbation so that every face and element mode is exercised
    y0 = ref.truncate(y0 + 0.01 * np.random.default_rng(9).standard_normal(y0.size))
    dt = ref.timestep(0.5)
    n = 6
    y = ref.rk4(y0, 0.0, dt, n)
    for prob in (fused, generic):
        integ = T.wave_integrator(prob, y0, 0.0, n * dt, dt)
        integ.solve()
        assert np.linalg.norm(integ.u - y) <= 1e-13 * np.linalg.norm(y)


@pytest.mark.parametrize("case", FUSED[:3], ids=_id)
def test_fused_f32c_matches_oracle(case):
    """FP32c fused stages (and the FP32c generic path) against the oracle's FP64 RK4."""
    forest = T.wave_forest(case["D"], case["P"], case["E"], roots=case["roots"],
                           reflecting=case.get("reflecting", False))
    p32 = T.WaveProblem(forest, T.Element(case["P"]), precision="f32c")
    g32 = T.WaveProblem(forest, T.Element(case["P"]), precision="f32c", generic=True)
    assert p32.info()["fused"]
    ref = O.Problem.brick(case["D"], case["P"], case["E"], roots=case["roots"],
                          reflecting=case.get("reflecting", False))
    y0 = ref.truncate(ref.fill_exact(1 if case.get("reflecting") else 0, 0.0, (1.0, 1)))
    dt = ref.timestep(0.5)
    n = 6
    y = ref.rk4(y0, 0.0, dt, n)
    for prob in (p32, g32):
        integ = T.wave_integrator(prob, y0, 0.0, n * dt, dt)
        integ.solve()
        assert np.linalg.norm(integ.u - y) <= 1e-5 * np.linalg.norm(y)


FUSED_REFINED = [
    dict(D=3, P=3, E=2, roots=4, regions=[((0.25, 0.75),) * 3]),
    dict(D=3, P=2, E=2, roots=4, reflecting=True, regions=[((0.0, 0.5),) * 3]),
    dict(D=3, P=4, E=4, roots=2, regions=[((0.5, 1.0),) * 3]),
    dict(D=3, P=3, E=2, roots=8, reflecting=True, regions=O.nested_cubes(3, 1 / 8, 3)),
]


@pytest.mark.parametrize("case", FUSED_REFINED, ids=_id)
def test_fused_refined_matches_oracle(case):
    """The ghosted fused stages (coarse-fine faces: face ghosts formed by two
    face-sized launches from the face buffers) against the oracle's RK4 and the
    generic exchange path."""
    prob, ref = _pair(case["D"], case["P"], case["E"], case["roots"],
                      reflecting=case.get("reflecting", False), regions=case["regions"])
    forest = T.wave_forest(case["D"], case["P"], case["E"], roots=case["roots"],
                           reflecting=case.get("reflecting", False), refined=list(case["regions"]))
    generic = T.WaveProblem(forest, T.Element(case["P"]), generic=True)
    assert prob.info()["fused"] and not generic.info()["fused"]
    refl = case.get("reflecting", False)
    y0 = ref.truncate(ref.fill_exact(1 if refl else 0, 0.0, (1.0, 2 if refl else 1)))
    y0 = ref.truncate(y0 + 0.01 * np.random.default_rng(11).standard_normal(y0.size))
    dt = ref.timestep(0.5)
    n = 5
    y = ref.rk4(y0, 0.0, dt, n)
    for p in (prob, generic):
        integ = T.wave_integrator(p, y0, 0.0, n * dt, dt)
        integ.solve()
        assert np.linalg.norm(integ.u - y) <= 1e-13 * np.linalg.norm(y)


@pytest.mark.parametrize("case", [FUSED[0], FUSED_REFINED[0]], ids=_id)
def test_fused_step_graph_replays_match_eager_steps(case):
    """n steps in one call replay the captured step graph; one step per call runs
    the launches eagerly.  Several calls with eager work and synchronization in
    between reuse the graph; every result equals the eager one bitwise."""
    forest = T.wave_forest(case["D"], case["P"], case["E"], roots=case["roots"],
                           reflecting=case.get("reflecting", False),
                           refined=list(case["regions"]) if case.get("regions") else False)
    a = T.WaveProblem(forest, T.Element(case["P"]))
    b = T.WaveProblem(forest, T.Element(case["P"]))
    assert a.info()["fused"]
    u0 = T.initial_state(a, T.wave_exact(3, 1.0, 1, 0.0))
    dt = T.oldg_timestep(a, 0.5)
    ia = T.wave_integrator(a, u0, 0.0, 9 * dt, dt)
    ib = T.wave_integrator(b, u0, 0.0, 9 * dt, dt)
    for n in (3, 2, 4):
        ia.step(n)                      # graph replays
        for _ in range(n):
            ib.step(1)                  # eager launches
        T.wave_energy(b, ib.u)          # other work on the device between calls
        assert np.array_equal(ia.u, ib.u), n


def test_fused_large_forests_each_first_problem():
    """Large forests of small blocks (one wave per workgroup, several
    workgroups per CU), each forest's first problem and a rebuilt one, one step
    against the oracle.  Guards the host staging of tables and states: staged
    in cacheable host memory, a fresh upload was at times read as the previous
    upload's contents by the first problem of a new forest.  The uploaded state
    must come back bit for bit before the step, and a mismatch after it names
    the blocks and variables it lies in."""
    for P, E, roots, regions, refl in [(3, 2, 11, (), False),
                                       (3, 2, 8, O.nested_cubes(3, 1 / 8, 3), True)]:
        forest = T.wave_forest(3, P, E, roots=roots, reflecting=refl,
                               refined=list(regions) if regions else False)
        ref = O.Problem.brick(3, P, E, roots=roots, reflecting=refl, regions=regions)
        y0 = ref.truncate(ref.fill_exact(1 if refl else 0, 0.0, (1.0, 2 if refl else 1)))
        dt = ref.timestep(0.5)
        y = ref.rk4(y0, 0.0, dt, 1)
        for rep in range(2):
            prob = T.WaveProblem(forest, T.Element(P))
            assert prob.info()["fused"]
            integ = T.wave_integrator(prob, y0, 0.0, dt, dt)
            assert np.array_equal(integ.u, y0), ("upload", roots, rep)
            integ.solve()
            got = integ.u
            err = np.linalg.norm(got - y)
            if not err <= 1e-13 * np.linalg.norm(y):
                # where it lies, whether a second download of the same state agrees,
                # whether the wrong blocks hold the state before the step, and
                # whether the same step again on the same problem is right
                bad = np.abs(got - y).reshape(prob.nblocks, 2, -1) > 1e-10 * np.max(np.abs(y))
                where = np.argwhere(bad.any(axis=2))
                redl = integ.u
                err_redl = np.linalg.norm(redl - y) / np.linalg.norm(y)
                blocks = np.unique(where[:, 0])
                g3 = got.reshape(prob.nblocks, 2, -1)[blocks]
                holds_y0 = bool(np.array_equal(g3, y0.reshape(prob.nblocks, 2, -1)[blocks]))
                again = T.wave_integrator(prob, y0, 0.0, dt, dt)
                again.solve()
                err2 = np.linalg.norm(again.u - y) / np.linalg.norm(y)
                raise AssertionError(
                    f"roots {roots} rep {rep}: relative error {err / np.linalg.norm(y):.3e}, "
                    f"{len(where)} (block, variable) pairs differ, first {where[:8].tolist()}; "
                    f"downloaded again: {err_redl:.3e}; wrong blocks hold y0: {holds_y0}; "
                    f"the same step run again on this problem: {err2:.3e}")
            prob.close()


def _ball(P=3, E=2, roots=(2, 2), precision="f64", truncate=True):
    from twoldg.geometry import sevenpatch_forest
    forest, geo = sevenpatch_forest(P, E, roots=roots)
    wave = T.PlaneWave((1.1, -0.7, 0.5), 0.3)
    prob = T.WaveProblem(forest, T.Element(P), truncate=truncate, precision=precision,
                         geometry=geo, boundary=wave)
    ref = O.Problem.ball(P, E, roots=roots, truncate=truncate)
    return prob, ref


@pytest.mark.parametrize("P", [2, 3])
def test_ball_rhs_truncate_energy_match_oracle(P, rough):
    """Full metric: Cartesian gradient, frame-weighted flux, glue ghosts and the
    Dirichlet sphere at t (evolution.jl:136-220, 1110-1127); per-element
    sqrt(g)-orthogonal truncation (evolution.jl:384-425); sqrt(g)-weighted energy."""
    prob, ref = _ball(P=P)
    u = rough(prob.state_len, seed=11)
    du = np.empty_like(u)
    T.wave_rhs(du, u, prob, 0.37)
    want = ref.rhs(u, 0.37)
    assert np.max(np.abs(du - want)) <= 1e-12 * np.max(np.abs(want))
    got = T.truncate_state(u.copy(), prob)
    want = ref.truncate(u)
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    assert T.wave_energy(prob, u, t=0.37) == pytest.approx(ref.energy(u, 0.37), rel=1e-12)
    l2 = T.wave_norm(prob, u)
    assert l2 == pytest.approx(ref.norms(u)[0], rel=1e-13)


@pytest.mark.parametrize("P", [2, 3])
def test_refined_ball_rhs_matches_oracle(P, rough):
    """The seven-patch ball with the middle of its cube refined (roots (4, 2)):
    coarse-fine faces on a full metric, whose Cartesian gradient ghosts are
    interpolated and restricted from the face-compact gradient, against the
    oracle's own refined ball: right-hand side at t = 0.37 and energy."""
    from twoldg.geometry import sevenpatch_forest
    forest, geo = sevenpatch_forest(P, 2, roots=(4, 2), refined=True)
    wave = T.PlaneWave((1.1, -0.7, 0.5), 0.3)
    prob = T.WaveProblem(forest, T.Element(P), geometry=geo, boundary=wave)
    ref = O.Problem.ball(P, 2, roots=(4, 2), regions=[((1.0, 3.0),) * 3])
    u = rough(prob.state_len, seed=17)
    du = np.empty_like(u)
    T.wave_rhs(du, u, prob, 0.37)
    want = ref.rhs(u, 0.37)
    assert np.max(np.abs(du - want)) <= 1e-12 * np.max(np.abs(want))
    assert T.wave_energy(prob, u, 0.37) == pytest.approx(ref.energy(u, 0.37), rel=1e-13)
    prob.close()


def test_ball_truncation_is_projection(rough):
    """sqrt(g)-orthogonal projection: idempotent (sevenpatch_tests.jl:128-158)."""
    prob, _ = _ball()
    u = rough(prob.state_len, seed=12)
    once = T.truncate_state(u.copy(), prob)
    twice = T.truncate_state(once.copy(), prob)
    assert np.max(np.abs(twice - once)) <= 1e-12 * np.max(np.abs(once))
    nn = prob.nn
    assert np.array_equal(once.reshape(-1, 2, nn)[:, 0], u.reshape(-1, 2, nn)[:, 0])


def test_sevenpatch_errors_published_and_oracle():
    """sevenpatch_errors(P=3, E=2, roots=(2,2)), k=(1.1,-0.7,0.5), phi=0.3, t=1:
    u 8.83e-3, v 1.14e-2, 1 - E1/E0 = 0.035 (CODE.md:4251), and the oracle's run."""
    r = T.sevenpatch_errors(3, 2, roots=(2, 2))
    o = O.sevenpatch_errors(3, 2, roots=(2, 2))
    assert r["nsteps"] == o["nsteps"]
    assert r["l2u"] == pytest.approx(8.83e-3, rel=5e-3)
    assert r["l2v"] == pytest.approx(1.14e-2, rel=5e-3)
    assert 1 - r["energy"] / r["energy0"] == pytest.approx(0.035, abs=5e-4)
    assert r["l2"] == pytest.approx(o["l2"], rel=SQRT_EPS)
    assert r["energy"] == pytest.approx(o["energy"], rel=SQRT_EPS)
    u = r["integrator"].u
    assert np.linalg.norm(u - o["y"]) <= SQRT_EPS * np.linalg.norm(o["y"])


# The seven-patch convergence table (CODE.md:4248-4251, bin/sevenpatch.jl: roots
# (2s, 2s) at scale s, E = 2, k = (1.1, -0.7, 0.5), phi = 0.3, cfl 1/2, t = 1):
# u's L2 error, v's, and 1 - E1/E0 at s = 1, 2, 3, printed to three digits.
SEVENPATCH_TABLE = {
    (2, True): ((6.07e-2, 1.63e-2, 7.34e-3), (7.81e-2, 2.25e-2, 1.04e-2), (0.016, 0.030, 0.033)),
    (2, False): ((4.97e-2, 3.87e-3, 9.09e-4), (9.33e-2, 1.45e-2, 6.44e-3), (0.032, 0.035, 0.036)),
    (3, True): ((8.83e-3, 1.15e-3, 3.36e-4), (1.14e-2, 1.38e-3, 3.84e-4), (0.035, 0.036, 0.036)),
    (3, False): ((2.32e-3, 2.38e-4, 7.09e-5), (6.85e-3, 9.03e-4, 2.50e-4), (0.035, 0.036, 0.036)),
}


@pytest.mark.parametrize("P,truncate", list(SEVENPATCH_TABLE),
                         ids=["P2-trunc", "P2-plain", "P3-trunc", "P3-plain"])
def test_sevenpatch_convergence_table_published(P, truncate):
    """Every row of the published seven-patch table at its three scales (up to
    1512 blocks and 1200 steps at s = 3), to the printed digits: the full metric,
    the glues, the Dirichlet sphere and both truncation settings converge as the
    reference's do."""
    us, vs, dEs = SEVENPATCH_TABLE[(P, truncate)]
    for s in (1, 2, 3):
        r = T.sevenpatch_errors(P, 2, roots=(2 * s, 2 * s), truncate=truncate)
        assert r["l2u"] == pytest.approx(us[s - 1], rel=5e-3), (s, r["l2u"])
        assert r["l2v"] == pytest.approx(vs[s - 1], rel=5e-3), (s, r["l2v"])
        assert 1 - r["energy"] / r["energy0"] == pytest.approx(dEs[s - 1], abs=5e-4), s
        r["problem"].close()


def test_sevenpatch_f32c_matches_oracle():
    """FP32 stencil with FP64-built coefficients stored as FP32 (the seven-patch
    ball's FP32 form) against the oracle's FP64 run to t = 1: state and l2 error."""
    r32 = T.sevenpatch_errors(3, 2, roots=(2, 2), precision="f32c")
    o = O.sevenpatch_errors(3, 2, roots=(2, 2))
    assert r32["l2"] == pytest.approx(o["l2"], rel=1e-4)
    u = r32["integrator"].u
    assert np.linalg.norm(u - o["y"]) <= 1e-4 * np.linalg.norm(o["y"])


def test_batch_matches_oracle_per_problem():
    """Batched uneven problems (orders 2..6, roots 4/8/16, Gaussian pulses), one
    launch per stage over the work list, graph-replayed steps: every problem
    equals the oracle's own RK4 run of it."""
    from twoldg.batch import gaussian_pulses, initial_states
    probs = gaussian_pulses(n=12, seed=7, roots=(2, 4))
    for prec, tol in (("f64", 1e-12), ("f32c", 2e-5)):
        batch = T.Batch([p[:3] for p in probs], precision=prec)
        u0s = initial_states(probs)
        for i, u0 in enumerate(u0s):
            batch.upload(i, u0)
        n = 5
        batch.rk4(2)
        batch.rk4(n - 2)
        info = batch.info()
        assert set(info["tiles"]) == {p[3]["P"] for p in probs}
        for i, p in enumerate(probs):
            ref = O.Problem.brick(2, p[3]["P"], 2, roots=p[3]["roots"])
            c, w = p[3]["c"], p[3]["w"]
            y0 = ref.truncate(ref.fill_exact(3, 0.0, (c[0], c[1], w)))
            assert np.max(np.abs(y0 - u0s[i])) <= 1e-15
            y = ref.rk4(y0, 0.0, p[2], n)
            got = batch.download(i)
            assert np.linalg.norm(got - y) <= tol * np.linalg.norm(y), (prec, i)
        batch.close()


def test_batch_every_block_shape_and_face_kind_matches_oracle():
    """One batch over the block shapes a batch takes (E = 2 and 4, orders 1 to 7),
    periodic and reflecting boxes, and a single-root box whose blocks are their
    own neighbours across both faces: every problem equals the oracle's RK4 run."""
    from twoldg.drivers import gaussian_pulse
    from twoldg.batch import initial_states
    cases = [(1, 4, 2, False), (2, 4, 2, True), (3, 4, 1, False), (7, 2, 2, False),
             (7, 2, 1, True), (1, 2, 3, True), (4, 4, 1, True), (5, 2, 3, False)]
    rng = np.random.default_rng(11)
    probs = []
    for P, E, roots, refl in cases:
        forest = T.wave_forest(2, P, E, roots=roots, reflecting=refl)
        c = rng.uniform(0.3, 0.7, size=2)
        w = float(rng.uniform(0.08, 0.2))
        dt = 0.5 * forest.minimum_spacing() / (P + 1)
        probs.append((forest, gaussian_pulse(c, w), dt, dict(P=P, E=E, roots=roots, refl=refl,
                                                             c=c, w=w)))
    u0s = initial_states(probs)
    n = 4
    want = []
    for i, p in enumerate(probs):
        m = p[3]
        ref = O.Problem.brick(2, m["P"], m["E"], roots=m["roots"], reflecting=m["refl"])
        y0 = ref.truncate(ref.fill_exact(3, 0.0, (m["c"][0], m["c"][1], m["w"])))
        assert np.max(np.abs(y0 - u0s[i])) <= 1e-15
        want.append(ref.rk4(y0, 0.0, p[2], n))
    for prec, tol in (("f64", 1e-12), ("f32c", 2e-5)):
        batch = T.Batch([p[:3] for p in probs], precision=prec)
        for i, u0 in enumerate(u0s):
            batch.upload(i, u0)
        batch.rk4(n)
        for i, y in enumerate(want):
            got = batch.download(i)
            err = np.linalg.norm(got - y) / np.linalg.norm(y)
            assert err <= tol, (prec, cases[i], err)
        batch.close()


def test_golden_fixtures():
    """The HIP path against the committed fixtures (tests/golden): RHS of a
    seeded state in 2D and on the ball (Dirichlet at t = 0.25), the ball's
    per-element truncation, and the drivers' results."""
    import json
    import os
    from twoldg.geometry import sevenpatch_forest
    G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    f = np.load(os.path.join(G, "rhs_2d_p4.npz"))
    prob = T.WaveProblem(T.wave_forest(2, 4, 2, roots=2), T.Element(4))
    du = np.empty_like(f["y"])
    T.wave_rhs(du, f["y"], prob, 0.0)
    assert np.max(np.abs(du - f["du"])) <= 1e-12 * np.max(np.abs(f["du"]))
    f = np.load(os.path.join(G, "rhs_ball_p1.npz"))
    forest, geo = sevenpatch_forest(1, 2, roots=(2, 2))
    pb = T.WaveProblem(forest, T.Element(1), geometry=geo,
                       boundary=T.PlaneWave((1.1, -0.7, 0.5), 0.3))
    du = np.empty_like(f["y"])
    T.wave_rhs(du, f["y"], pb, 0.25)
    assert np.max(np.abs(du - f["du"])) <= 1e-12 * np.max(np.abs(f["du"]))
    tr = T.truncate_state(f["y"].copy(), pb)
    assert np.max(np.abs(tr - f["ytrunc"])) <= 1e-12 * np.max(np.abs(f["ytrunc"]))
    res = json.load(open(os.path.join(G, "results.json")))
    runs = {"wave_d2_p4_r8": T.wave_errors(2, 4, 2, roots=8),
            "reflect_d2_p2_r4": T.reflect_errors(2, 2, 2, roots=4),
            "twolevel_d2_p3_r4": T.wave_errors(2, 3, 2, roots=4, refined=True),
            "sevenpatch_p3_s1": T.sevenpatch_errors(3, 2, roots=(2, 2))}
    for name, r in runs.items():
        want = res[name]
        assert r["nsteps"] == want["nsteps"], name
        for k in ("l2", "l2u", "l2v", "linf", "energy0", "energy"):
            assert r[k] == pytest.approx(want[k], rel=SQRT_EPS), (name, k)


@pytest.mark.parametrize("which", ["ball", "2d"])
def test_generic_step_graph_replays_match_eager_steps(which):
    """The generic path (the seven-patch ball with its Dirichlet sphere, and a 2D
    forest) replays a captured RK4 step; the timed run of the same steps is
    eager.  With the same call pattern both give the same state bit for bit:
    the Dirichlet data read each step's time from the device clock."""
    if which == "ball":
        from twoldg.geometry import sevenpatch_forest
        forest, geo = sevenpatch_forest(3, 2, roots=(2, 2))
        wave = T.PlaneWave((1.1, -0.7, 0.5), 0.3)
        mk = lambda: T.WaveProblem(forest, T.Element(3), geometry=geo, boundary=wave)
        f0 = T.sevenpatch_exact(wave, 0.0)
    else:  # the generic stages (a 2D uniform problem otherwise runs the batch stage)
        forest = T.wave_forest(2, 4, 2, roots=4)
        mk = lambda: T.WaveProblem(forest, T.Element(4), generic=True)
        f0 = T.wave_exact(2, 1.0, 1, 0.0)
    a, b = mk(), mk()
    assert not a.info()["fused"]
    u0 = T.initial_state(a, f0)
    dt = T.oldg_timestep(a, 0.5)
    ia = T.wave_integrator(a, u0, 0.0, 20 * dt, dt)
    ib = T.wave_integrator(b, u0, 0.0, 20 * dt, dt)
    for n in (5, 3, 7):
        ia.step(n)                      # graph replays
        ib.timed(n)                     # eager launches, event-timed
        T.wave_energy(b, ib.u, t=ib.t)  # other work on the device between calls
        assert np.array_equal(ia.u, ib.u), n


@pytest.mark.parametrize("P", [2, 3, 4])
def test_annulus_rhs_truncate_energy_norms_match_oracle(P, rough):
    """The diagonal metric on the four-patch annulus (flux A_d D_d u, divergence
    over sqrt(g), identity glues in theta, walls in r, per-line truncation pairs,
    sqrt(g)-weighted energy and norms; evolution.jl:67-110, 556-575, 1472-1512)
    against the oracle on a scrambled state."""
    forest, geo = T.annulus_forest(P, 2, roots=(2, 2))
    prob = T.WaveProblem(forest, T.Element(P), geometry=geo)
    assert not prob.info()["fused"]
    ref = O.Problem.annulus(P, 2, roots=(2, 2))
    u = rough(prob.state_len, seed=13)
    du = np.empty_like(u)
    T.wave_rhs(du, u, prob, 0.0)
    want = ref.rhs(u)
    assert np.max(np.abs(du - want)) <= 1e-12 * np.max(np.abs(want))
    got = T.truncate_state(u.copy(), prob)
    want = ref.truncate(u)
    assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))
    assert T.wave_energy(prob, u) == pytest.approx(ref.energy(u), rel=1e-13)
    assert T.wave_norm(prob, u) == pytest.approx(ref.norms(u)[0], rel=1e-14)
    assert T.wave_norm(prob, u, math.inf) == ref.norms(u)[1]
    assert T.oldg_timestep(prob, 0.5) == ref.timestep(0.5)


def test_annulus_errors_match_oracle_and_published():
    """annulus_errors(P=3, E=2, roots=(2s, s)), m=2, n=1, to a quarter turn: the
    published u 5.06e-3, 6.30e-4 and v 3.35e-3, 4.23e-4 at s = 1, 2 (CODE.md:4194),
    the oracle's run to sqrt(eps) (state, l2, energy), and FP32c against it."""
    k = T.annulus_wavenumber(2, 1, 1.0, 2.0)
    for s, eu, ev in ((1, 5.06e-3, 3.35e-3), (2, 6.30e-4, 4.23e-4)):
        r = T.annulus_errors(3, 2, roots=(2 * s, s), m=2)
        o = O.annulus_errors(3, 2, roots=(2 * s, s), m=2, k=k)
        assert r["k"] == k and r["nsteps"] == o["nsteps"]
        assert float(f"{r['l2u']:.3g}") == eu and float(f"{r['l2v']:.3g}") == ev
        assert r["l2"] == pytest.approx(o["l2"], rel=SQRT_EPS)
        assert r["energy"] == pytest.approx(o["energy"], rel=SQRT_EPS)
        assert r["energy0"] == pytest.approx(o["energy0"], rel=1e-13)
        u = r["integrator"].u
        assert np.linalg.norm(u - o["y"]) <= SQRT_EPS * np.linalg.norm(o["y"])
        r32 = T.annulus_errors(3, 2, roots=(2 * s, s), m=2, precision="f32c")
        assert r32["l2"] == pytest.approx(o["l2"], rel=1e-3)
        u32 = r32["integrator"].u
        assert np.linalg.norm(u32 - o["y"]) <= 1e-5 * np.linalg.norm(o["y"])
