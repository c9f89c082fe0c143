"""Invariants the reference's own tests hold its kernels to, on the library.

* No kernel reads a working-array entry it has not written
  (refined_tests.jl:145-161): the reference poisons its working arrays with NaN
  before a right-hand side.  Here every working array of a problem (stage
  states, ghost rows, gradient sets, face buffers and ghost slots, reduction
  partials) is filled with NaN between two otherwise identical calls, and the
  second must equal the first bit for bit, on every path: the generic stages
  (flat, refined, reflecting), the fused 3D stages in each precision (uniform
  and with coarse-fine faces), the batch stage of a 2D problem, the seven-patch
  ball (Dirichlet sphere), the thirteen-patch ball (radiative sphere, factored
  truncation) and the annulus (diagonal metric).
* With the identity map the full-metric path is the flat path, bit for bit
  (sevenpatch_tests.jl:37-61): S = I and sqrt(g) = 1 on a periodic brick make
  every frame product a product with an exact one or zero.  The full metric is
  three-dimensional here, so D = 3 only.
* A block whose frame is a constant diagonal gives the same right-hand side on
  the flat-block branch and on the general one (sevenpatch_tests.jl:526-559):
  the seven-patch ball's central cube, flagged flat and then not.
"""
import numpy as np
import pytest

import twoldg as T
from twoldg.geometry import SevenPatchBall, sevenpatch_forest

pytestmark = pytest.mark.gpu


def _state(n, seed):
    return np.random.default_rng(seed).standard_normal(n)


# -- problems on every path -------------------------------------------------------------
def _flat(D, P, E, roots, precision="f64", reflecting=False, regions=(), generic=False):
    forest = T.wave_forest(D, P, E, roots=roots, reflecting=reflecting,
                           refined=list(regions) if regions else False)
    return T.WaveProblem(forest, T.Element(P), precision=precision, generic=generic)


def _ball7(P=2):
    forest, geo = sevenpatch_forest(P, 2, roots=(2, 2))
    return T.WaveProblem(forest, T.Element(P), geometry=geo,
                         boundary=T.PlaneWave((1.1, -0.7, 0.5), 0.3))


def _ball13(P=2):
    forest, geo = T.thirteenpatch_forest(P, 2, roots=(2, 1, 2), Lc=1, Rm=2, R1=3.5)
    return T.WaveProblem(forest, T.Element(P), geometry=geo, boundary=T.RadiativeBoundary())


def _annulus(P=3):
    forest, geo = T.annulus_forest(P, 2, roots=(2, 2))
    return T.WaveProblem(forest, T.Element(P), geometry=geo)


HALF2 = [((0.25, 0.75),) * 2]
HALF3 = [((0.25, 0.75),) * 3]
# three levels of nested cubes, finest half-width 1/8
NESTED = [((0.25, 0.75),) * 3, ((0.375, 0.625),) * 3]

PATHS = {
    "2d-refined": (lambda: _flat(2, 3, 2, 4, regions=HALF2), "generic"),
    "2d-refined-reflecting": (lambda: _flat(2, 2, 2, 4, reflecting=True, regions=HALF2),
                              "generic"),
    "3d-refined-generic": (lambda: _flat(3, 2, 2, 4, regions=HALF3, generic=True), "generic"),
    "3d-uniform-f64": (lambda: _flat(3, 3, 2, 2), "fused3d"),
    "3d-uniform-f32c": (lambda: _flat(3, 3, 2, 2, precision="f32c"), "fused3d"),
    "3d-uniform-split": (lambda: _flat(3, 3, 2, 2, precision="split"), "fused3d"),
    "3d-nested-f64": (lambda: _flat(3, 3, 2, 4, reflecting=True, regions=NESTED), "fused3d"),
    "3d-nested-f32c": (lambda: _flat(3, 3, 2, 4, precision="f32c", reflecting=True,
                                     regions=NESTED), "fused3d"),
    "3d-nested-split": (lambda: _flat(3, 3, 2, 4, precision="split", reflecting=True,
                                      regions=NESTED), "fused3d"),
    "3d-nested-r8-split": (lambda: _flat(3, 3, 2, 8, precision="split", reflecting=True,
                                         regions=NESTED), "fused3d"),
    "2d-batch-f64": (lambda: _flat(2, 4, 2, 4), "batch2d"),
    "2d-batch-f32c": (lambda: _flat(2, 4, 2, 4, precision="f32c"), "batch2d"),
    "ball7": (_ball7, "generic"),
    "ball13-radiative": (_ball13, "generic"),
    "annulus": (_annulus, "generic"),
}


@pytest.mark.parametrize("name", list(PATHS))
def test_poisoned_working_arrays_change_nothing(name):
    """Every entry point returns the same bits after its problem's working
    arrays were filled with NaN: the right-hand side, the truncation, the
    energy and norms, and three RK4 steps of an integrator (which replay the
    step graph of the first run)."""
    make, stages = PATHS[name]
    p = make()
    assert p.info()["stages"] == stages
    u = _state(p.state_len, 21)
    t = 0.3

    def pure():
        du = np.empty_like(u)
        T.wave_rhs(du, u, p, t)
        tr = T.truncate_state(u.copy(), p)
        return du, tr, T.wave_energy(p, u, t), T.wave_norm(p, u), T.wave_norm(p, u, np.inf)

    clean = pure()
    p.poison()
    poisoned = pure()
    assert np.all(np.isfinite(poisoned[0]))
    for a, b in zip(clean, poisoned):
        assert np.array_equal(a, b)

    dt = T.oldg_timestep(p, 0.5)
    y0 = T.truncate_state(u.copy(), p) * 1e-2

    def steps(poison):
        integ = T.wave_integrator(p, y0, 0.0, 10 * dt, dt)
        integ.step(1)
        if poison:
            p.poison()
        integ.step(3)
        return integ.u

    clean = steps(False)
    poisoned = steps(True)
    assert np.all(np.isfinite(poisoned))
    assert np.array_equal(clean, poisoned)
    p.close()


# -- the full metric's flat cases --------------------------------------------------------
class IdentityFrame:
    """The identity map on a brick as a full geometry: sqrt(g) = 1, S[d] = e_d at
    every node, no outer faces (the reference's test geometry of the same name,
    sevenpatch_tests.jl:19-26).  `flat` flags every block for the flat-block
    branch, or none."""
    kind = "full"
    outer_faces = False

    def __init__(self, flat=True):
        self.flag = 1 if flat else 0

    def frame(self, patch, xi):
        n = xi.shape[1]
        return np.ones(n), np.broadcast_to(np.eye(3), (n, 3, 3)).copy()

    def _element_blocks(self, forest):
        return np.arange(forest.nblocks)

    def metric(self, forest, el):
        nb, nn = forest.nblocks, forest.N ** 3
        coef = np.zeros((nb, 10, nn))
        coef[:, 0] = 1.0
        for d in range(3):
            coef[:, 1 + 3 * d + d] = 1.0
        flat = np.full(nb, self.flag, dtype=np.uint8)
        return dict(coef=coef, flat=flat, hspeed=float(np.min(forest.spacings())))

    # the per-element sqrt(g)-orthogonal projection tables, J = 1 on every block
    truncation_tables = SevenPatchBall.truncation_tables


@pytest.mark.parametrize("flat", [True, False], ids=["flat-blocks", "general"])
@pytest.mark.parametrize("P", [2, 3])
def test_identity_map_full_metric_is_the_flat_path(P, flat):
    """On a periodic brick with sqrt(g) = 1 and S = I the Cartesian gradient holds
    D_i u in component i and the flux F^d that component plus exact zeros, so
    the right-hand side is the flat one's to the bit, on the flat-block branch
    and on the general one.  The per-element truncation is then the element's
    tensor projector built another way: equal to rounding."""
    forest = T.wave_forest(3, P, 2, roots=2)
    el = T.Element(P)
    pf = T.WaveProblem(forest, el)
    pc = T.WaveProblem(forest, el, geometry=IdentityFrame(flat))
    assert pc.info()["stages"] == "generic"
    u = _state(pf.state_len, 0)
    duf, duc = np.empty_like(u), np.empty_like(u)
    T.wave_rhs(duf, u, pf, 0.0)
    T.wave_rhs(duc, u, pc, 0.0)
    assert np.array_equal(duf, duc)
    assert T.wave_energy(pc, u) == pytest.approx(T.wave_energy(pf, u), rel=1e-14)
    a, b = T.truncate_state(u.copy(), pc), T.truncate_state(u.copy(), pf)
    assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))
    pf.close()
    pc.close()


class _NotFlat:
    """A full geometry with every block sent through the general path."""

    def __init__(self, geo):
        self._geo = geo

    def __getattr__(self, name):
        return getattr(self._geo, name)

    def metric(self, forest, el):
        m = dict(self._geo.metric(forest, el))
        m["flat"] = np.zeros_like(m["flat"])
        return m


@pytest.mark.parametrize("P", [2, 3])
def test_constant_diagonal_frame_flat_branch_is_the_general_one(P):
    """The seven-patch ball's central cube (a constant diagonal frame) on the
    flat-block branch and on the general one: the same right-hand side to the
    bit; the shells' frames vary and are never flagged."""
    forest, geo = sevenpatch_forest(P, 2, roots=(4, 1), refined=True)
    wave = T.PlaneWave((1.1, -0.7, 0.5), 0.3)
    el = T.Element(P)
    p = T.WaveProblem(forest, el, geometry=geo, boundary=wave)
    cube = np.array([lf[0] == 0 for lf in forest.leaves])
    assert np.array_equal(p.flat.astype(bool), cube)
    q = T.WaveProblem(forest, el, geometry=_NotFlat(geo), boundary=wave)
    assert not q.flat.any()
    u = _state(p.state_len, 1)
    a, b = np.empty_like(u), np.empty_like(u)
    T.wave_rhs(a, u, p, 0.0)
    T.wave_rhs(b, u, q, 0.0)
    assert np.array_equal(a, b)
    p.close()
    q.close()
