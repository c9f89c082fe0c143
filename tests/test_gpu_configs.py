"""GPU parity at each benchmarked workload's own size and form, against the oracle.

Every figure the benchmark reports runs on a workload that one of these tests
holds to the oracle at that size:

* 3D periodic forest, P=4, E=4, roots=16 (4096 blocks of 20^3, 320^3 nodes):
  two RK4 steps of the fused stages, FP64 and FP32c;
* the reflecting box with nested cubes (P=3, E=2, roots=8, 1408 blocks) in
  FP32c, every refined forest of the ghosted fused stage;
* the 512 batched 2D problems drawn with seed 20261203 (orders 2..6, roots
  4/8/16), FP64 and FP32c, one problem per (P, roots) pair against the oracle;
* the seven-patch ball at scale 2 (roots (4,4), 448 blocks) with the plane wave
  k = 2 pi (1,1,1)/sqrt(3), phi = 0, to t = 1, FP64 and FP32c.

Tolerances: the reference's device tests (device_tests.jl:70-74) hold a device
run to sqrt(eps) of the host's; FP64 here is held to 1e-13 relative (the runs
differ only in summation order), FP32c to the rounding of FP32 storage.
"""
import math

import numpy as np
import pytest

import oracle as O
import twoldg as T

pytestmark = pytest.mark.gpu

SQRT_EPS = math.sqrt(np.finfo(np.float64).eps)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def f32c_floor(dt):
    """The relative distance FP32 storage of u alone puts between an FP32c run and
    an FP64 one after a few steps at cfl 1/2, and no less than 1e-5: u's rounding
    (2^-24 relative) turned into v by dt times the Laplacian, whose largest
    eigenvalue is omega_max^2 with dt omega_max = 0.49 (RK4's imaginary-axis
    limit 2 sqrt(2) times cfl 0.5 over the 3D stability limit 2.9, which the
    oracle's power iteration pins), so 2^-24 * 0.49^2 / dt."""
    return max(1e-5, 2.0 ** -24 * 0.49 ** 2 / dt)


@pytest.mark.timeout(600)
def test_periodic_320_cubed_two_steps_match_oracle():
    """The 320^3 periodic forest: the oracle's state after two RK4 steps, FP64
    to 1e-13 and FP32c to the FP32 storage floor (relative L2 over the whole
    state); at dt = 3.125e-4 that floor is 4.6e-5."""
    D, P, E, roots = 3, 4, 4, 16
    ref = O.Problem.brick(D, P, E, roots=roots)
    y0 = ref.truncate(ref.fill_exact(0, 0.0, (1.0, 1)))
    dt = ref.timestep(0.5)
    y = ref.rk4(y0, 0.0, dt, 2)
    forest = T.wave_forest(D, P, E, roots=roots)
    for prec, tol in (("f64", 1e-13), ("f32c", f32c_floor(dt))):
        prob = T.WaveProblem(forest, T.Element(P), precision=prec)
        assert prob.info()["fused"]
        assert T.oldg_timestep(prob, 0.5) == dt
        integ = T.wave_integrator(prob, y0, 0.0, 2 * dt, dt)
        integ.solve()
        err = _rel(integ.u, y)
        print(f"320^3 {prec}: relative L2 distance from the oracle after 2 steps {err:.3e}",
              flush=True)
        assert err <= tol, (prec, err)
        prob.close()


NESTED = [
    dict(D=3, P=3, E=2, roots=4, regions=[((0.25, 0.75),) * 3]),
    dict(D=3, P=2, E=2, roots=4, reflecting=True, regions=[((0.0, 0.5),) * 3]),
    dict(D=3, P=4, E=4, roots=2, regions=[((0.5, 1.0),) * 3]),
    dict(D=3, P=3, E=2, roots=8, reflecting=True, regions=O.nested_cubes(3, 1 / 8, 3)),
]


@pytest.mark.parametrize("case", NESTED, ids=["ref1", "reflref1", "p4ref1", "nested3"])
def test_refined_f32c_matches_oracle(case):
    """FP32c on every refined forest of the ghosted fused stage, the nested cubes
    (1408 blocks, reflecting walls, cos mode m=2) included: 10 RK4 steps
    against the oracle's FP64 run, to the FP32 storage floor at the finest
    level's dt (2.9e-5 on the nested cubes)."""
    refl = case.get("reflecting", False)
    forest = T.wave_forest(case["D"], case["P"], case["E"], roots=case["roots"], reflecting=refl,
                           refined=list(case["regions"]))
    prob = T.WaveProblem(forest, T.Element(case["P"]), precision="f32c")
    assert prob.info()["fused"]
    ref = O.Problem.brick(case["D"], case["P"], case["E"], roots=case["roots"], reflecting=refl,
                          regions=case["regions"])
    y0 = ref.truncate(ref.fill_exact(1 if refl else 0, 0.0, (1.0, 2 if refl else 1)))
    dt = ref.timestep(0.5)
    n = 10
    y = ref.rk4(y0, 0.0, dt, n)
    integ = T.wave_integrator(prob, y0, 0.0, n * dt, dt)
    integ.solve()
    err = _rel(integ.u, y)
    print(f"refined f32c: relative L2 distance from the oracle {err:.3e}", flush=True)
    assert err <= f32c_floor(dt)
    assert integ.energy() == pytest.approx(ref.energy(y), rel=1e-5)


@pytest.mark.timeout(600)
def test_batch_512_seed_20261203_matches_oracle():
    """The benchmarked batch (512 problems, seed 20261203, orders 2..6, roots
    4/8/16, E=2, Gaussian pulses): every problem advanced together, then the
    first problem of each (P, roots) pair against the oracle's run of it; FP64
    to 1e-12 and FP32c to 2e-5 after 5 steps."""
    from twoldg.batch import gaussian_pulses, initial_states
    probs = gaussian_pulses()
    assert len(probs) == 512
    pairs = {}
    for i, p in enumerate(probs):
        pairs.setdefault((p[3]["P"], p[3]["roots"]), i)
    assert len(pairs) == 15  # every (P, roots) pair is drawn
    u0s = initial_states(probs)
    n = 5
    want = {}
    for (P, roots), i in pairs.items():
        ref = O.Problem.brick(2, P, 2, roots=roots)
        c, w = probs[i][3]["c"], probs[i][3]["w"]
        y0 = ref.truncate(ref.fill_exact(3, 0.0, (c[0], c[1], w)))
        assert np.max(np.abs(y0 - u0s[i])) <= 1e-15
        want[i] = ref.rk4(y0, 0.0, probs[i][2], n)
    for prec, tol in (("f64", 1e-12), ("f32c", 2e-5)):
        batch = T.Batch([p[:3] for p in probs], precision=prec)
        for i, u0 in enumerate(u0s):
            batch.upload(i, u0)
        batch.rk4(n)
        for i, y in want.items():
            err = _rel(batch.download(i), y)
            assert err <= tol, (prec, i, probs[i][3]["P"], probs[i][3]["roots"], err)
        batch.close()


@pytest.mark.timeout(600)
def test_ball_scale2_baseline_wave_matches_oracle():
    """The seven-patch ball at scale 2 (448 blocks) with the plane wave
    k = 2 pi (1,1,1)/sqrt(3), phi = 0, Dirichlet data on the sphere, to t = 1:
    FP64 and FP32c against the oracle's run (state, l2 error, energy)."""
    k = tuple(2 * math.pi / math.sqrt(3) for _ in range(3))
    o = O.sevenpatch_errors(3, 2, roots=(4, 4), k=k, phi=0.0)
    for prec, tol in (("f64", SQRT_EPS), ("f32c", 1e-4)):
        r = T.sevenpatch_errors(3, 2, roots=(4, 4), k=k, phi=0.0, precision=prec)
        assert r["nsteps"] == o["nsteps"]
        u = r["integrator"].u
        err = _rel(u, o["y"])
        print(f"ball s=2 {prec}: state vs oracle {err:.3e}, l2 {r['l2']:.6e} "
              f"(oracle {o['l2']:.6e})", flush=True)
        assert err <= tol, (prec, err)
        assert r["l2"] == pytest.approx(o["l2"], rel=tol)
        assert r["energy"] == pytest.approx(o["energy"], rel=tol)
        r["problem"].close()


def test_periodic2d_f32c_matches_oracle():
    """The 80^2 periodic forest (P=4, E=2, roots=8) in FP32c against the oracle's
    FP64 run: the L2 error against the exact mode equals the oracle's to 1%, and
    the states agree to FP32 rounding."""
    r = T.wave_errors(2, 4, 2, roots=8, precision="f32c")
    o = O.wave_errors(2, 4, 2, roots=8)
    assert r["l2"] == pytest.approx(o["l2"], rel=1e-2)
    assert _rel(r["integrator"].u, o["y"]) <= 1e-5
