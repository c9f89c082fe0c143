"""The split form of the fused stages: u and its stencil in FP64, v in FP32 with
its step update added to an FP32 (hi, lo) pair.

Against the oracle's FP64 RK4 on the fused forests (uniform and refined), the
split form's error is that of FP32 rounding in v's stage values and
accumulator, far below the compensated FP32 form's, whose floor is the FP32
stencil acting on an FP32 u.
"""
import numpy as np
import pytest

import oracle as O
import twoldg as T

pytestmark = pytest.mark.gpu

CASES = [
    dict(D=3, P=3, E=2, roots=2),
    dict(D=3, P=2, E=2, roots=4, reflecting=True),
    dict(D=3, P=4, E=4, roots=2),                       # the 320^3 workload's 20^3 blocks
    dict(D=3, P=8, E=2, roots=2),
    dict(D=3, P=5, E=4, roots=2, reflecting=True),
    dict(D=3, P=3, E=2, roots=4, regions=[((0.25, 0.75),) * 3]),
    dict(D=3, P=3, E=2, roots=8, reflecting=True, regions=O.nested_cubes(3, 1 / 8, 3)),
]


def _id(c):
    return "D{D}P{P}E{E}r{roots}".format(**c) + ("refl" if c.get("reflecting") else "") + \
        ("ref" if c.get("regions") else "")


@pytest.mark.parametrize("case", CASES, ids=_id)
def test_split_matches_oracle(case):
    refl = case.get("reflecting", False)
    regions = case.get("regions", ())
    forest = T.wave_forest(case["D"], case["P"], case["E"], roots=case["roots"], reflecting=refl,
                           refined=list(regions) if regions else False)
    ps = T.WaveProblem(forest, T.Element(case["P"]), precision="split")
    p32 = T.WaveProblem(forest, T.Element(case["P"]), precision="f32c")
    assert ps.info()["fused"]
    ref = O.Problem.brick(case["D"], case["P"], case["E"], roots=case["roots"],
                          reflecting=refl, regions=regions)
    y0 = ref.truncate(ref.fill_exact(1 if refl else 0, 0.0, (1.0, 2 if refl else 1)))
    y0 = ref.truncate(y0 + 0.01 * np.random.default_rng(5).standard_normal(y0.size))
    dt = ref.timestep(0.5)
    n = 6
    y = ref.rk4(y0, 0.0, dt, n)
    errs = {}
    for name, prob in (("split", ps), ("f32c", p32)):
        integ = T.wave_integrator(prob, y0, 0.0, n * dt, dt)
        integ.solve()
        errs[name] = np.linalg.norm(integ.u - y) / np.linalg.norm(y)
    assert errs["split"] <= 2e-7
    assert errs["split"] < 0.2 * errs["f32c"]
    # the state crosses the boundary as FP64: an upload downloads unchanged
    ps.upload(y0)
    assert np.array_equal(ps.download(), y0)
    for p in (ps, p32):
        p.close()


def test_split_steps_one_at_a_time_and_replayed():
    forest = T.wave_forest(3, 3, 2, roots=2)
    ps = T.WaveProblem(forest, T.Element(3), precision="split")
    u0 = T.initial_state(ps, T.wave_exact(3, 1.0, 1, 0.0))
    dt = T.oldg_timestep(ps, 0.5)
    a = T.wave_integrator(ps, u0, 0.0, 8 * dt, dt)
    a.solve()
    ua = a.u
    b = T.wave_integrator(ps, u0, 0.0, 8 * dt, dt)
    for _ in range(8):
        b.step(1)
    ub = b.u
    # each call splits v into its FP32 pair and joins it again: 2^-48 apart
    assert np.abs(ua - ub).max() <= 1e-13 * np.abs(ua).max()
    # the energy after the steps is FP64's (RK4 loses the same amount) to the
    # FP32 rounding of v
    p64 = T.WaveProblem(forest, T.Element(3))
    c = T.wave_integrator(p64, u0, 0.0, 8 * dt, dt)
    c.solve()
    e0, e1, e64 = T.wave_energy(ps, u0), T.wave_energy(ps, ua), T.wave_energy(p64, c.u)
    assert abs(e1 - e64) <= 1e-7 * e0
    ps.close()
    p64.close()


def test_split_refuses_what_has_no_fused_stage():
    for forest in (T.wave_forest(2, 3, 2, roots=2), T.wave_forest(3, 3, 6, roots=1)):
        with pytest.raises(ValueError, match="split form"):
            T.WaveProblem(forest, T.Element(3), precision="split")
