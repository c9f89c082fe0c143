"""The thirteen-patch ball, its factored truncation, and the radiative boundary
(geometry.jl:555-749, evolution.jl:296-380, 904-984, 1129-1299,
sevenpatch.jl:131-236).

The CPU tests pin the host map and the oracle: the frame against the map's
finite differences and its stated structure (thirteenpatch_tests.jl:47-119),
the oracle's frame (forward-mode dual numbers) against the host's closed form,
the energy's rate across every glue and through the radiative sphere (a
quadratic energy's rate along the right-hand side, by an exact central
difference), and the reference's published runs: the plane wave on the
thirteen-patch ball (CODE.md:4517-4535) and the outgoing pulse through the
radiative sphere on both balls (CODE.md:4268-4280, 4540-4545).  The GPU tests
compare the library with the oracle, check the factored truncation against the
per-element projection of the same map, and repeat the rate and published
checks on the device.
"""
import math

import numpy as np
import pytest

import oracle as O
import twoldg as T
from twoldg import _lib

gpu = pytest.mark.gpu


# -- helpers ---------------------------------------------------------------------------
def pulse(x, t, w=0.5):
    """(f(t - r) - f(t + r))/r and its time derivative, f(s) = exp(-s^2/w^2), with
    the r -> 0 limit -2 f'(t), -2 f''(t) below r = w/1000."""
    r = np.sqrt((x ** 2).sum(axis=0))

    def f(s):
        return np.exp(-(s / w) ** 2)

    def df(s):
        return -2 * s / w ** 2 * f(s)

    def d2f(s):
        return (4 * s * s / w ** 4 - 2 / w ** 2) * f(s)

    far = r > w / 1000
    rr = np.where(far, r, 1.0)
    u = np.where(far, (f(t - rr) - f(t + rr)) / rr, -2 * df(t))
    v = np.where(far, (df(t - rr) - df(t + rr)) / rr, -2 * d2f(t))
    return u, v


def oracle_state(o, fn):
    x = o.coords()
    nb, _, nn = x.shape
    u, v = fn(np.moveaxis(x, 1, 0).reshape(3, -1))
    y = np.empty((nb, 2, nn))
    y[:, 0] = u.reshape(nb, nn)
    y[:, 1] = v.reshape(nb, nn)
    return y.reshape(-1)


def outer_faces(forest, geo):
    """(block, k, side) of every block face on the outer sphere."""
    first = 7 if isinstance(geo, T.ThirteenPatchBall) else 1
    out = []
    for b, lf in enumerate(forest.leaves):
        p = lf[0]
        if p < first:
            continue
        k, s = [(0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1)][p - first]
        side = 1 if s == 1 else 0
        n = forest.dims(p, lf[1])
        if lf[2 + k] == (n[k] - 1 if side else 0):
            out.append((b, k, side))
    return out


def radiative_rate(forest, geo, coef, y, kappa):
    """-kappa sum w |S[k]| v^2 over the outer faces' nodes: the energy's rate,
    the boundary's own energy included."""
    N, P1 = forest.N, forest.P + 1
    wq = T.Element(forest.P).weights
    h = forest.spacings()
    ya = y.reshape(forest.nblocks, 2, N, N, N)        # [b, var, z, y, x]
    rate = 0.0
    for b, k, side in outer_faces(forest, geo):
        i = N - 1 if side else 0
        idx = [slice(None)] * 3
        idx[2 - k] = i
        v = ya[b, 1][tuple(idx)]
        S = coef[b, 1 + 3 * k:4 + 3 * k].reshape(3, N, N, N)[(slice(None),) + tuple(idx)]
        S = np.sqrt((S ** 2).sum(axis=0))
        w = np.ones((N, N))
        tdims = [e for e in range(3) if e != k]      # plane axes (slow, fast) = (t1, t0)
        w *= (h[b] * P1 / 2 * wq[np.arange(N) % P1])[None, :]
        w *= (h[b] * P1 / 2 * wq[np.arange(N) % P1])[:, None]
        rate -= kappa * float(np.sum(w * S * v * v))
    return rate


def energy_rate(energy, rhs, y):
    """The rate of a quadratic energy along the right-hand side,
    (E(y + e f) - E(y - e f)) / 2e (exact for a quadratic form), and the
    Cauchy-Schwarz bound 2 sqrt(E(y) E(f)) of its terms."""
    f = rhs(y)
    ey, ef = energy(y), energy(f)
    eps = math.sqrt(ey / ef)
    return (energy(y + eps * f) - energy(y - eps * f)) / (2 * eps), 2 * math.sqrt(ey * ef)


def zero_outer_layer(forest, geo, y):
    """u and v zero in the outer shells' outermost elements, so that nothing
    crosses the outer sphere under zero data (thirteenpatch_tests.jl:30-45)."""
    N, P1 = forest.N, forest.P + 1
    ya = y.reshape(forest.nblocks, 2, N, N, N).copy()
    for b, k, side in outer_faces(forest, geo):
        idx = [slice(None)] * 3
        idx[2 - k] = slice(N - P1, N) if side else slice(0, P1)
        ya[b][(slice(None),) + tuple(idx)] = 0.0
    return ya.reshape(-1)


# ---------------------------------------------------------------------------
# The host map, pinned.
# ---------------------------------------------------------------------------
def test_thirteen_frame_is_the_maps():
    forest, geo = T.thirteenpatch_forest(2, 2, roots=(2, 1, 2), Lc=1, Rm=2, R1=3.5)
    assert len(forest.patches) == 13
    x = np.array([[0.3], [0.7], [0.9]])
    eps = 1e-6
    for q in range(13):
        J, S = geo.frame(q, x)
        M = np.empty((3, 3))
        for e in range(3):
            xp, xm = x.copy(), x.copy()
            xp[e] += eps
            xm[e] -= eps
            M[:, e] = (geo.position(q, xp) - geo.position(q, xm))[:, 0] / (2 * eps)
        Sfd = abs(np.linalg.det(M)) * np.linalg.inv(M)
        assert np.abs(S[0] - Sfd).max() < 1e-8 * max(1, J[0])
        assert J[0] > 0
    from twoldg.geometry import SHELLS
    for i, (k, s) in enumerate(SHELLS):
        qi, qo = i + 1, i + 7

        def at(xk):
            v = np.full((3, 1), 0.6)
            v[k] = xk
            return v

        inner, middle = at(0.0 if s == 1 else 1.0), at(1.0 if s == 1 else 0.0)
        want = np.full(3, -0.4)
        want[k] = s
        assert np.array_equal(geo.position(qi, inner)[:, 0], want)
        assert math.isclose(np.linalg.norm(geo.position(qi, middle)), 2, rel_tol=1e-15)
        start, outer = at(0.0 if s == 1 else 2.0), at(2.0 if s == 1 else 0.0)
        assert np.allclose(geo.position(qo, start), geo.position(qi, middle), rtol=0,
                           atol=4e-16)
        assert math.isclose(np.linalg.norm(geo.position(qo, outer)), 3.5, rel_tol=1e-15)
        for xi in ((0.3, 0.7, 0.9), (1.9, 0.1, 1.2)):
            xv = np.array(xi).reshape(3, 1)
            X = geo.position(qo, xv)[:, 0]
            rho = xi[k] / 2 if s == 1 else 1 - xi[k] / 2
            assert math.isclose(np.linalg.norm(X), 2 + 1.5 * rho, rel_tol=1e-14)
            c = np.array([s if j == k else 2 * xi[j] / 2 - 1 for j in range(3)], dtype=float)
            assert np.allclose(X, np.linalg.norm(X) * c / np.linalg.norm(c), rtol=1e-14)
            J, S = geo.frame(qo, xv)
            Sk = S[0, k]
            assert np.linalg.norm(np.cross(Sk, X)) < 1e-13 * np.linalg.norm(Sk) * np.linalg.norm(X)
            for d in range(3):
                if d != k:
                    assert abs(S[0, d] @ X) < 1e-13 * np.linalg.norm(S[0, d]) * np.linalg.norm(X)
    assert geo.outer_radius() == 3.5


def test_thirteen_refusals():
    with pytest.raises(ValueError, match="odd"):
        T.thirteenpatch_forest(2, 3)
    with pytest.raises(ValueError, match="√3 Lc"):
        T.thirteenpatch_forest(2, 2, Lc=1, Rm=1.5, R1=3)
    with pytest.raises(ValueError, match="must exceed the middle radius"):
        T.thirteenpatch_forest(2, 2, Rm=2, R1=2)
    with pytest.raises(ValueError, match="(n, m, m′)"):
        T.thirteenpatch_forest(2, 2, roots=(2, 2))
    with pytest.raises(ValueError, match="positive number of roots"):
        T.thirteenpatch_forest(2, 2, roots=(2, 1, 0))
    with pytest.raises(ValueError, match="reaches the cube's faces"):
        T.thirteenpatch_forest(2, 2, roots=(4, 1, 1), refined=((0, 2), (0, 2), (0, 2)))
    with pytest.raises(ValueError, match="κ = -1"):
        T.RadiativeBoundary(kappa=-1.0)
    with pytest.raises(ValueError, match="not graded"):
        T.ball_forest(2, 2, roots=(2, 1, 1), Rm=2.0, grading=0.5)


def test_outer_shell_volume_is_the_product_of_its_factors():
    # thirteenpatch_tests.jl:97-119: r^2 r' times (2/n)^2 q^(-3/2), the
    # angular factor the same function on all six shells
    forest, geo = T.thirteenpatch_forest(2, 2, roots=(3, 1, 2), Lc=1, Rm=2, R1=3.5)
    assert all(geo.factored_volume(q) == -1 for q in range(7))
    for xi in ((0.3, 0.7, 0.9), (2.9, 0.1, 1.2), (1.5, 1.5, 1.5)):
        xv = np.array(xi).reshape(3, 1)
        for q in range(7, 13):
            k = geo.factored_volume(q)
            assert k == [0, 0, 1, 1, 2, 2][q - 7]
            t = [xi[j] for j in range(3) if j != k]
            J = geo.frame(q, xv)[0][0]
            assert math.isclose(geo.radial_volume(xi[k], q) * geo.angular_volume(*t), J,
                                rel_tol=1e-14)


def test_thirteen_glue_and_sphere_faces():
    forest, geo = T.thirteenpatch_forest(2, 2, roots=(2, 1, 1))
    faces = forest.faces()
    kinds = [faces[i].kind for i in range(forest.nblocks * 6)]
    assert kinds.count(_lib.FACE_DIRICHLET) == 6 * 2 ** 2
    assert len(outer_faces(forest, geo)) == 6 * 2 ** 2
    assert all(forest.leaves[b][0] >= 7 for b, _, _ in outer_faces(forest, geo))


# ---------------------------------------------------------------------------
# The oracle, pinned.
# ---------------------------------------------------------------------------
def test_oracle_thirteen_forest_and_frame_are_the_hosts():
    el = T.Element(2)
    forest, geo = T.thirteenpatch_forest(2, 2, roots=(2, 1, 2), Lc=1, Rm=2, R1=3.5)
    o = O.Problem.ball_ext(2, 2, roots=(2, 1, 2), Lc=1, Rm=2, R1=3.5)
    assert [(b[0], b[1], b[2]) for b in o.blocks()] == \
        [(lf[0], lf[1], tuple(lf[2:5])) for lf in forest.leaves]
    assert np.array_equal(o.coords(), geo.physical(forest, forest.node_logical(el.offsets)))
    c, flat = o.coef()
    m = geo.metric(forest, el)
    assert np.abs(c - m["coef"]).max() < 1e-13 * np.abs(m["coef"]).max()
    assert np.array_equal(flat, m["flat"].astype(bool))


@pytest.mark.parametrize("P", [2, 3])
@pytest.mark.parametrize("truncate", [False, True])
def test_oracle_energy_invariant_across_thirteen_glues(P, truncate, rough):
    # thirteenpatch_tests.jl:136-152: zero data on the sphere and a state that
    # vanishes in the outermost elements: nothing crosses, so the energy's rate
    # is rounding
    forest, geo = T.thirteenpatch_forest(P, 2, roots=(2, 1, 2))
    o = O.Problem.ball_ext(P, 2, roots=(2, 1, 2), truncate=truncate, k=(0, 0, 0), phi=0.0)
    y = zero_outer_layer(forest, geo, rough(o.size(), 2))
    y = o.truncate(y) if truncate else y
    rate, scale = energy_rate(lambda z: o.energy(z), lambda z: o.rhs(z), y)
    assert scale > 1
    assert abs(rate) < 1e-13 * scale


@pytest.mark.parametrize("thirteen", [False, True])
@pytest.mark.parametrize("P", [2, 3])
def test_oracle_radiative_energy_rate(thirteen, P, rough):
    # sevenpatch_tests.jl:314-341, thirteenpatch_tests.jl:154-172: the energy,
    # the boundary's own included, changes at -kappa sum w |S| v^2 alone
    roots = (2, 1, 1) if thirteen else (2, 2)
    kw = dict(Rm=2.0) if thirteen else {}
    forest, geo = T.ball_forest(P, 2, roots=roots, **kw)
    for kappa, ir in ((1.0, True), (1.0, False), (0.5, True), (0.0, False)):
        for truncate in (False, True):
            o = O.Problem.ball_ext(P, 2, roots=roots, Rm=2.0, R1=3.0, truncate=truncate,
                                   radiative=True, kappa=kappa, invR=1 / 3 if ir else 0.0)
            y = rough(o.size(), 3)
            y = o.truncate(y) if truncate else y
            rate, scale = energy_rate(lambda z: o.energy(z), lambda z: o.rhs(z), y)
            want = radiative_rate(forest, geo, o.coef()[0], y, kappa)
            assert scale > 1
            assert abs(rate - want) < 1e-13 * scale
            if kappa > 0:
                assert want < -1e-8 * scale   # far above the tolerance of the rate check


@pytest.mark.parametrize("P,steps,u,v", [(2, 239, 6.28e-2, 8.09e-2), (3, 424, 9.37e-3, 1.24e-2)])
def test_oracle_thirteen_plane_wave_published(P, steps, u, v):
    # CODE.md:4521-4531: the plane wave on the thirteen-patch ball Lc 1, Rm 2,
    # R1 3, roots (2, 1, 1), to t = 1
    r = O.sevenpatch_errors(P, 2, roots=(2, 1, 1), Rm=2.0, tend=1.0)
    assert r["nsteps"] == steps
    assert math.isclose(r["l2u"], u, rel_tol=5e-3)
    assert math.isclose(r["l2v"], v, rel_tol=5e-3)


@pytest.mark.parametrize("roots,E1E0,l2", [((2, 2), 3.5e-3, 2.5e-2), ((2, 1, 1), 5.0e-3, 2.3e-2)])
def test_oracle_outgoing_pulse_published(roots, E1E0, l2):
    # CODE.md:4268-4280 (seven-patch) and 4540-4545 (thirteen-patch, Rm = 2):
    # a width-1/2 pulse under Sommerfeld's condition at P = 3 to t = 5
    o = O.Problem.ball_ext(3, 2, roots=roots, Rm=2.0, R1=3.0, radiative=True, kappa=1.0,
                           invR=1 / 3)
    y0 = o.truncate(oracle_state(o, lambda x: pulse(x, 0.0)))
    n = math.ceil(5.0 / o.timestep(0.5))
    y = o.rk4(y0, 0.0, 5.0 / n, n)
    assert math.isclose(o.energy(y) / o.energy(y0), E1E0, rel_tol=0.03)
    err = y - oracle_state(o, lambda x: pulse(x, 5.0))
    assert math.isclose(o.norms(err)[0], l2, rel_tol=0.03)


# ---------------------------------------------------------------------------
# The device.
# ---------------------------------------------------------------------------
def _problem(P, roots, Rm=None, boundary=None, truncate=True, precision="f64", factored=True):
    if Rm is None:
        forest, geo = T.sevenpatch_forest(P, 2, roots=roots)
    else:
        forest, geo = T.thirteenpatch_forest(P, 2, roots=roots, Rm=Rm, factored=factored)
    return T.WaveProblem(forest, T.Element(P), geometry=geo, truncate=truncate,
                         boundary=boundary, precision=precision), forest, geo


@gpu
@pytest.mark.parametrize("P", [2, 3])
def test_gpu_thirteen_rhs_truncate_energy_match_oracle(P, rough):
    wave = T.PlaneWave((1.1, -0.7, 0.5), 0.3)
    p, forest, geo = _problem(P, (2, 1, 2), Rm=2.0, boundary=wave)
    o = O.Problem.ball_ext(P, 2, roots=(2, 1, 2), Rm=2.0, R1=3.0)
    y = rough(p.state_len, 4)
    for t in (0.0, 0.37):
        du = np.empty_like(y)
        T.wave_rhs(du, y, p, t)
        want = o.rhs(y, t)
        assert np.abs(du - want).max() < 1e-12 * np.abs(want).max()
    yt = T.truncate_state(y.copy(), p)
    want = o.truncate(y)
    assert np.abs(yt - want).max() < 1e-13 * np.abs(want).max()
    assert math.isclose(T.wave_energy(p, yt, t=0.2), o.energy(yt, 0.2), rel_tol=1e-13)
    p.close()


@gpu
def test_gpu_thirteen_e4_radiative_matches_oracle(rough):
    """The benchmark ball's element layout (P = 3, E = 4: 16^3 blocks, factored
    truncation on the outer shells) under the radiative sphere with u/R: the
    right-hand side, the truncation and the energy against the oracle."""
    P, E, roots = 3, 4, (2, 1, 1)
    forest, geo = T.thirteenpatch_forest(P, E, roots=roots, Rm=2.0)
    p = T.WaveProblem(forest, T.Element(P), geometry=geo, boundary=T.RadiativeBoundary())
    o = O.Problem.ball_ext(P, E, roots=roots, Rm=2.0, R1=3.0, radiative=True, kappa=1.0,
                           invR=1 / 3)
    y = rough(p.state_len, 8)
    yt = T.truncate_state(y.copy(), p)
    want = o.truncate(y)
    assert np.abs(yt - want).max() < 1e-13 * np.abs(want).max()
    du = np.empty_like(yt)
    T.wave_rhs(du, yt, p, 0.0)
    want = o.rhs(yt)
    assert np.abs(du - want).max() < 1e-12 * np.abs(want).max()
    assert math.isclose(T.wave_energy(p, yt), o.energy(yt), rel_tol=1e-13)
    p.close()


@gpu
@pytest.mark.parametrize("P", [2, 3])
def test_gpu_factored_truncation_is_the_per_element_projection(P, rough):
    # thirteenpatch_tests.jl:121-165: the radial pair times the angular
    # projection is the sqrt(g)-orthogonal projection onto Q_{P-1}; 48 outer
    # blocks share four angular tables
    zero = T.ZeroData()
    p, forest, geo = _problem(P, (2, 1, 2), Rm=2.0, boundary=zero)
    pe, _, _ = _problem(P, (2, 1, 2), Rm=2.0, boundary=zero, factored=False)
    t = p._tables
    outer = np.array([lf[0] >= 7 for lf in forest.leaves])
    assert outer.sum() == 48
    assert np.array_equal(t["fslot"] >= 0, outer)
    assert np.array_equal(t["slot"] >= 0, (pe._tables["slot"] >= 0) & ~outer)
    assert t["nangular"] == 4 and t["ntables"] == 24
    a, b = rough(p.state_len, 1), rough(p.state_len, 6)
    Pa, Pb = T.truncate_state(a.copy(), p), T.truncate_state(b.copy(), p)
    assert np.abs(Pa - T.truncate_state(a.copy(), pe)).max() < 1e-14 * np.abs(Pa).max() * 10
    nb, nn = forest.nblocks, forest.N ** 3
    c, _ = O.Problem.ball_ext(P, 2, roots=(2, 1, 2), Rm=2.0).coef()
    el = T.Element(P)
    w1 = el.weights[np.arange(forest.N) % (P + 1)] * (P + 1) / 2
    w = np.multiply.outer(np.multiply.outer(w1, w1), w1).reshape(-1)
    W = np.outer(forest.spacings() ** 3, w) * c[:, 0]

    def inner(x, y):
        return float(np.sum(W * x.reshape(nb, 2, nn)[:, 1] * y.reshape(nb, 2, nn)[:, 1]))

    scale = math.sqrt(inner(a, a) * inner(b, b))
    assert abs(inner(Pa, b) - inner(a, Pb)) < 1e-14 * scale * 10
    assert np.allclose(T.truncate_state(Pa.copy(), p), Pa, rtol=0, atol=1e-13 * np.abs(Pa).max())
    assert np.array_equal(Pa.reshape(nb, 2, nn)[:, 0], a.reshape(nb, 2, nn)[:, 0])
    assert inner(Pa, Pa) <= inner(a, a)
    p.close()
    pe.close()


@gpu
@pytest.mark.parametrize("P", [2, 3])
@pytest.mark.parametrize("truncate", [False, True])
def test_gpu_energy_invariant_across_thirteen_glues(P, truncate, rough):
    p, forest, geo = _problem(P, (2, 1, 2), Rm=2.0, boundary=T.ZeroData(), truncate=truncate)
    y = zero_outer_layer(forest, geo, rough(p.state_len, 2))
    y = T.truncate_state(y, p)

    def rhs(z):
        out = np.empty_like(z)
        T.wave_rhs(out, z, p, 0.0)
        return out

    rate, scale = energy_rate(lambda z: T.wave_energy(p, z, t=0.0), rhs, y)
    assert scale > 1
    assert abs(rate) < 1e-13 * scale
    p.close()


@gpu
@pytest.mark.parametrize("thirteen", [False, True])
@pytest.mark.parametrize("P", [2, 3])
def test_gpu_radiative_energy_rate_and_oracle(thirteen, P, rough):
    roots = (2, 1, 1) if thirteen else (2, 2)
    for kappa, ir in ((1.0, True), (1.0, False), (0.5, True), (0.0, False)):
        for truncate in (False, True):
            p, forest, geo = _problem(P, roots, Rm=2.0 if thirteen else None,
                                      boundary=T.RadiativeBoundary(kappa, ir),
                                      truncate=truncate)
            assert p.rad_invR == (1 / 3 if ir else 0.0)
            y = T.truncate_state(rough(p.state_len, 3), p)

            def rhs(z):
                out = np.empty_like(z)
                T.wave_rhs(out, z, p, 0.0)
                return out

            rate, scale = energy_rate(lambda z: T.wave_energy(p, z), rhs, y)
            c = p.coef
            want = radiative_rate(forest, geo, c, y, kappa)
            assert abs(rate - want) < 1e-13 * scale
            if kappa > 0:
                assert want < -1e-8 * scale   # far above the tolerance of the rate check
            o = O.Problem.ball_ext(P, 2, roots=roots, Rm=2.0, R1=3.0, truncate=truncate,
                                   radiative=True, kappa=kappa, invR=1 / 3 if ir else 0.0)
            f = rhs(y)
            assert np.abs(f - o.rhs(y)).max() < 1e-12 * np.abs(f).max()
            assert math.isclose(T.wave_energy(p, y), o.energy(y), rel_tol=1e-13)
            p.close()


@gpu
def test_gpu_radiative_refusals():
    forest = T.wave_forest(3, 2, 2, roots=2)
    with pytest.raises(ValueError, match="full geometry carries"):
        T.WaveProblem(forest, boundary=T.RadiativeBoundary())


@gpu
@pytest.mark.parametrize("P,steps,u,v", [(2, 239, 6.28e-2, 8.09e-2), (3, 424, 9.37e-3, 1.24e-2)])
def test_gpu_thirteen_plane_wave_published_and_oracle(P, steps, u, v):
    r = T.sevenpatch_errors(P, 2, roots=(2, 1, 1), Rm=2.0, tend=1.0)
    assert r["nsteps"] == steps
    assert math.isclose(r["l2u"], u, rel_tol=5e-3)
    assert math.isclose(r["l2v"], v, rel_tol=5e-3)
    o = O.sevenpatch_errors(P, 2, roots=(2, 1, 1), Rm=2.0, tend=1.0)
    y = r["integrator"].u
    assert np.abs(y - o["y"]).max() < 1e-10 * np.abs(o["y"]).max()
    r["problem"].close()


@gpu
@pytest.mark.parametrize("P", [2, 3])
def test_gpu_thirteen_plane_wave_converges_at_order_P(P):
    rs = [T.sevenpatch_errors(P, 2, roots=(2 * s, s, s), Rm=2.0, tend=0.25) for s in (1, 2)]
    rate = T.convergence_rate([r["h"] for r in rs], [r["l2"] for r in rs])
    assert P - 0.3 < rate < P + 0.4
    for r in rs:
        r["problem"].close()


@gpu
@pytest.mark.parametrize("Rm,roots,E1E0,l2", [(None, (2, 2), 3.5e-3, 2.5e-2),
                                               (2.0, (2, 1, 1), 5.0e-3, 2.3e-2)])
def test_gpu_outgoing_pulse_published(Rm, roots, E1E0, l2):
    r = T.radiative_errors(3, 2, roots=roots, Rm=Rm, tend=5.0)
    assert math.isclose(r["energy"] / r["energy0"], E1E0, rel_tol=0.03)
    assert math.isclose(r["l2"], l2, rel_tol=0.03)
    r["problem"].close()
