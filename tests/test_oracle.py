"""Pins of the oracle (CPU): closed forms, invariants and the reference's
published tables, each chosen so that a plausible mistake (a dropped term, a
wrong sign, a transposed operand, a wrong coarse-fine operator) fails one.

References: element_tests.jl, evolution_tests.jl, refined_tests.jl,
sevenpatch_tests.jl, test/reference.jl and the tables in CODE.md.
"""
import math

import numpy as np
import pytest

import oracle as O
import twoldg as T


def legendre(P, x):
    return np.polynomial.legendre.legval(x, [0] * P + [1])


# --- element tables (element_tests.jl) ----------------------------------------
def test_paper_matrices():
    """The paper's printed matrices (element_tests.jl:38-56, Figs. 5-7)."""
    W2 = O.element(2)["stencil"] * 3
    assert np.allclose(W2, [[-3, 0, 4, -1, 0], [0, -1, 0, 1, 0], [0, 1, -4, 0, 3]], atol=1e-14)
    W3 = np.round(O.element(3)["stencil"] * 4, 2)
    assert np.array_equal(W3, [[-6, 0, 8.09, -3.09, 1, 0], [0, -1.62, 0, 2.24, -0.62, 0],
                               [0, 0.62, -2.24, 0, 1.62, 0], [0, -1, 3.09, -8.09, 0, 6]])
    W4 = np.round(O.element(4)["stencil"] * 5, 2)
    assert np.array_equal(W4, [[-10, 0, 13.51, -5.33, 2.82, -1, 0],
                               [0, -2.48, 0, 3.49, -1.53, 0.52, 0],
                               [0, 0.75, -2.67, 0, 2.67, -0.75, 0],
                               [0, -0.52, 1.53, -3.49, 0, 2.48, 0],
                               [0, 1, -2.82, 5.33, -13.51, 0, 10]])


@pytest.mark.parametrize("P", range(1, 9))
def test_gll_quadrature_and_structure(P):
    e = O.element(P)
    x, w, W = e["nodes"], e["weights"], e["stencil"]
    assert x[0] == -1 and x[-1] == 1 and np.all(x == -x[::-1]) and np.all(np.diff(x) > 0)
    assert e["offsets"][0] == 0 and e["offsets"][-1] == P + 1
    for k in range(2 * P):
        assert abs(np.sum(w * x ** k) - (2 / (k + 1) if k % 2 == 0 else 0)) < 1e-14
    assert abs(np.sum(w * x ** (2 * P)) - 2 / (2 * P + 1)) > 1e-14     # GLL, not Gauss
    assert np.all(W[1:, 0] == 0) and np.all(W[:-1, -1] == 0)
    assert W[0, 0] != 0 and W[-1, -1] != 0
    assert all(abs(W[i, i + 1]) < 1e-12 for i in range(P + 1))       # zero diagonal
    assert np.allclose(W, -W[::-1, ::-1], atol=1e-12)                   # odd under x -> -x


@pytest.mark.parametrize("P", range(1, 9))
def test_wide_derivative_exact_across_continuous_face(P):
    e = O.element(P)
    x, W = e["nodes"], e["stencil"]
    h = 2 / (P + 1)
    for k in range(P + 1):
        u = x ** k
        window = np.concatenate([[u[0]], u, [u[-1]]])
        du = k * x ** (k - 1) if k else 0 * x
        assert np.allclose(W @ window / h, du, atol=1e-11 * max(1, P * P))


def periodic_line(P, K, h):
    """The dense wide derivative, projector and weights on a periodic line of K
    elements (test/reference.jl:9-29)."""
    e = O.element(P)
    P1 = P + 1
    n = K * P1
    Dm = np.zeros((n, n))
    Pi = np.zeros((n, n))
    w = np.zeros(n)
    for q in range(K):
        for i in range(P1):
            r = q * P1 + i
            w[r] = h * P1 / 2 * e["weights"][i]
            for j in range(-1, P1 + 1):
                Dm[r, (q * P1 + j) % n] += e["stencil"][i, j + 1] / h
            for j in range(P1):
                Pi[r, q * P1 + j] = e["trunc"][i, j]
    return Dm, Pi, w


@pytest.mark.parametrize("P", range(1, 7))
@pytest.mark.parametrize("K", [1, 3])
def test_summation_by_parts_on_periodic_line(P, K):
    Dm, _, w = periodic_line(P, K, 0.1)
    WD = w[:, None] * Dm
    assert np.linalg.norm(WD + WD.T) <= 1e-12 * np.linalg.norm(WD)


@pytest.mark.parametrize("P", range(1, 9))
def test_truncation_projector(P):
    e = O.element(P)
    x, w, Pi = e["nodes"], e["weights"], e["trunc"]
    for k in range(P):
        assert np.allclose(Pi @ x ** k, x ** k, atol=1e-12)
    assert np.linalg.norm(Pi @ legendre(P, x)) < 1e-12
    assert np.allclose(Pi @ Pi, Pi, atol=1e-12)
    assert np.allclose(w[:, None] * Pi, (w[:, None] * Pi).T, atol=1e-12)
    assert np.allclose(Pi, np.eye(P + 1) - np.outer(e["topmode"], e["topdual"]), atol=1e-14)


# --- the RHS against a dense global operator (evolution_tests.jl:22-59) -------------
def global_index(forest, D, N):
    """Global node index of each owned node: block corner c*N plus the owned index."""
    n = forest.patches[0][0][0] * N
    idx = []
    I = np.stack(np.meshgrid(*[np.arange(N)] * D, indexing="ij"), 0)
    I = I.transpose(0, *reversed(range(1, D + 1))).reshape(D, -1)  # first dim fastest
    for (p, lev, *c) in forest.leaves:
        g = np.zeros(I.shape[1], dtype=np.int64)
        s = 1
        for d in range(D):
            g += (c[d] * N + I[d]) * s
            s *= n
        idx.append(g)
    return np.array(idx), n


@pytest.mark.parametrize("D,P,E,roots", [(1, 3, 2, 3), (2, 2, 2, 2), (2, 4, 2, 1), (3, 1, 2, 1)])
@pytest.mark.parametrize("truncate", [True, False])
def test_rhs_equals_dense_operator(D, P, E, roots, truncate):
    forest = T.wave_forest(D, P, E, roots=roots)
    ref = O.Problem.brick(D, P, E, roots=roots, truncate=truncate)
    N = E * (P + 1)
    idx, n = global_index(forest, D, N)
    h = 1.0 / n
    Dm, _, _ = periodic_line(P, n // (P + 1), h)
    I = np.eye(n)
    lap = 0
    for d in range(D):
        A = np.ones((1, 1))
        for e in reversed(range(D)):
            A = np.kron(A, Dm @ Dm if e == d else I)
        lap = lap + A
    rng = np.random.default_rng(2)
    y = rng.standard_normal(ref.size())
    du = ref.rhs(y)
    nn = N ** D
    u = np.zeros(n ** D)
    ys = y.reshape(-1, 2, nn)
    u[idx.reshape(-1)] = ys[:, 0].reshape(-1)
    want = (lap @ u)[idx.reshape(-1)].reshape(-1, nn)
    got = du.reshape(-1, 2, nn)
    assert np.allclose(got[:, 0], ys[:, 1], rtol=0, atol=0)
    assert np.max(np.abs(got[:, 1] - want)) <= 1e-12 * np.max(np.abs(want))


def test_truncation_equals_tensor_projector():
    """(Pi x Pi) per element (evolution_tests.jl:83-103), 2D."""
    P, E = 3, 2
    ref = O.Problem.brick(2, P, E, roots=1)
    Pi = O.element(P)["trunc"]
    N = E * (P + 1)
    y = np.random.default_rng(3).standard_normal(ref.size())
    got = ref.truncate(y).reshape(2, N, N)
    v = y.reshape(2, N, N)[1]
    want = np.zeros((N, N))
    for ey in range(E):
        for ex in range(E):
            blk = v[ey * (P + 1):(ey + 1) * (P + 1), ex * (P + 1):(ex + 1) * (P + 1)]
            want[ey * (P + 1):(ey + 1) * (P + 1), ex * (P + 1):(ex + 1) * (P + 1)] = Pi @ blk @ Pi.T
    assert np.max(np.abs(got[1] - want)) < 1e-13
    assert np.array_equal(got[0], y.reshape(2, N, N)[0])


# --- coarse-fine: skew-adjointness and dE/dt = 0 (refined_tests.jl:82-143) ----------
def node_weights(forest, P):
    el = T.Element(P)
    D, N = forest.D, forest.N
    h = forest.spacings()
    wi = el.weights[np.arange(N) % (P + 1)] * (P + 1) / 2
    w1 = np.ones(1)
    for d in range(D):
        w1 = np.kron(wi, w1)  # first dimension fastest
    return h[:, None] ** D * w1[None, :]


REFINED = [
    (2, 3, 2, 4, False, [((0.25, 0.75),) * 2]),
    (2, 2, 2, 4, True, [((0.0, 0.5),) * 2]),
    (3, 2, 2, 4, False, [((0.25, 0.75),) * 3]),
    (3, 3, 2, 8, True, O.nested_cubes(3, 1 / 8, 3)),
]


@pytest.mark.parametrize("case", REFINED)
def test_skew_adjoint_across_coarse_fine_faces(case):
    D, P, E, roots, refl, reg = case
    forest = T.wave_forest(D, P, E, roots=roots, reflecting=refl, refined=list(reg))
    ref = O.Problem.brick(D, P, E, roots=roots, reflecting=refl, regions=reg)
    w = node_weights(forest, P)
    nn = forest.N ** D
    rng = np.random.default_rng(4)
    y1 = rng.standard_normal(ref.size())
    u = y1.reshape(-1, 2, nn)[:, 0]
    Du = ref.gradients(y1).reshape(D, -1, nn)
    for d in range(D):
        # u exchanged as the state (even), g as the gradient's component d (odd)
        g = rng.standard_normal(u.shape)
        Dg = ref.divergence(g, d)
        a = np.sum(w * g * Du[d])
        b = np.sum(w * Dg * u)
        scale = (math.sqrt(np.sum(w * g * g) * np.sum(w * Du[d] ** 2)) +
                 math.sqrt(np.sum(w * Dg * Dg) * np.sum(w * u * u)))
        assert abs(a + b) < 1e-13 * scale


@pytest.mark.parametrize("case", REFINED)
def test_energy_rate_vanishes(case):
    """dE/dt = 0 for the semi-discrete system; E is quadratic, so the rate along
    f(y) is (E(y + f) - E(y - f)) / 2 exactly."""
    D, P, E, roots, refl, reg = case
    ref = O.Problem.brick(D, P, E, roots=roots, reflecting=refl, regions=reg)
    y = ref.truncate(np.random.default_rng(5).standard_normal(ref.size()))
    f = ref.rhs(y)
    rate = (ref.energy(y + f) - ref.energy(y - f)) / 2
    assert abs(rate) < 1e-12 * (ref.energy(y) + ref.energy(f))


def test_energy_rate_across_ball_glues():
    """Every glue of the seven-patch ball conserves energy (sevenpatch_tests.jl:
    182-226): zero Dirichlet data (k = 0), the state zero in the shells' outer
    elements."""
    from twoldg.geometry import sevenpatch_forest
    for P in (2, 3):
        for trunc in (False, True):
            forest, geo = sevenpatch_forest(P, 2, roots=(2, 2))
            ref = O.Problem.ball(P, 2, roots=(2, 2), truncate=trunc, k=(0, 0, 0), phi=0.0)
            N = forest.N
            nn = N ** 3
            xl = forest.node_logical(T.Element(P).offsets)
            y = np.random.default_rng(6).standard_normal(ref.size()).reshape(-1, 2, nn)
            he = (P + 1) / N
            for b, lf in enumerate(forest.leaves):
                if lf[0] == 0:
                    continue
                k, s = [(0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1)][lf[0] - 1]
                xk = xl[b, k]
                mask = xk >= 2 - he - 1e-12 if s == 1 else xk <= he + 1e-12
                y[b, :, mask] = 0
            y = ref.truncate(y.reshape(-1))
            f = ref.rhs(y)
            rate = (ref.energy(y + f) - ref.energy(y - f)) / 2
            assert abs(rate) < 1e-12 * (ref.energy(y) + ref.energy(f)), (P, trunc, rate)


# --- published tables ------------------------------------------------------------------
def test_wave_errors_2d_p4_published_table():
    """CODE.md:4142 (D=2, P=4, truncate): roots 2/4/8 u 1.46e-3 1.02e-4 6.61e-6,
    v 1.20e-2 9.31e-4 6.38e-5; CODE.md:4143 (no truncation) roots 8: 1.05e-7, 2.69e-5."""
    for roots, u, v in [(2, 1.46e-3, 1.20e-2), (4, 1.02e-4, 9.31e-4), (8, 6.61e-6, 6.38e-5)]:
        r = O.wave_errors(2, 4, 2, roots=roots)
        assert float(f"{r['l2u']:.3g}") == u and float(f"{r['l2v']:.3g}") == v
        assert r["energy"] <= r["energy0"]
    r = O.wave_errors(2, 4, 2, roots=8, truncate=False)
    assert float(f"{r['l2u']:.3g}") == 1.05e-7 and float(f"{r['l2v']:.3g}") == 2.69e-5


def test_reflecting_published_table():
    """CODE.md:4169 (D=2, P=2): 9.31e-2 2.64e-2 6.05e-3, E1/E0 = 1.0."""
    for roots, l2 in [(2, 9.31e-2), (4, 2.64e-2), (8, 6.05e-3)]:
        r = O.wave_errors(2, 2, 2, roots=roots, reflecting=True)
        assert float(f"{r['l2']:.3g}") == l2
        assert r["energy"] / r["energy0"] == pytest.approx(1.0, abs=1e-6)


def test_two_level_published_tables():
    """CODE.md:4626 and :4636 (D=2, P=3, middle half refined): L2 9.63e-3
    1.17e-3 1.45e-4, L-inf 4.83e-2 4.94e-3 9.65e-4 at roots 4, 8, 16."""
    mid = [((0.25, 0.75),) * 2]
    for roots, l2, li in [(4, 9.63e-3, 4.83e-2), (8, 1.17e-3, 4.94e-3), (16, 1.45e-4, 9.65e-4)]:
        r = O.wave_errors(2, 3, 2, roots=roots, regions=mid)
        assert float(f"{r['l2']:.3g}") == l2 and float(f"{r['linf']:.3g}") == li


def test_sevenpatch_published_table():
    """CODE.md:4251: P=3, truncated, s=1: u 8.83e-3, v 1.14e-2, 1 - E1/E0 = 0.035;
    no truncation: u 2.32e-3, v 6.85e-3."""
    r = O.sevenpatch_errors(3, 2, roots=(2, 2))
    assert float(f"{r['l2u']:.3g}") == 8.83e-3 and float(f"{r['l2v']:.3g}") == 1.14e-2
    assert 1 - r["energy"] / r["energy0"] == pytest.approx(0.035, abs=5e-4)
    r = O.sevenpatch_errors(3, 2, roots=(2, 2), truncate=False)
    assert float(f"{r['l2u']:.3g}") == 2.32e-3 and float(f"{r['l2v']:.3g}") == 6.85e-3


def test_rk4_stability_limit():
    """The RK4 limit cfl * sqrt(D) lies in (4.8, 5.8) (evolution_tests.jl:197-221,
    CODE.md:4725-4731): the spectral radius of the semi-discrete operator by
    power iteration, times dt, against RK4's imaginary-axis limit 2 sqrt(2)."""
    for D in (1, 2):
        P, E, roots = 3, 2, 2
        ref = O.Problem.brick(D, P, E, roots=roots)
        y = np.random.default_rng(8).standard_normal(ref.size())
        lam = 0
        for _ in range(300):
            z = ref.rhs(ref.rhs(y))          # L^2 has eigenvalues -omega^2
            lam = np.linalg.norm(z) / np.linalg.norm(y)
            y = z / np.linalg.norm(z)
        omega = math.sqrt(lam)
        h = ref.min_spacing()
        cfl_max = 2 * math.sqrt(2) / omega / (h / (P + 1))
        assert 4.8 < cfl_max * math.sqrt(D) < 5.8, cfl_max * math.sqrt(D)


def test_sevenpatch_published_table_scale2():
    """CODE.md:4251 at s=2 (roots (4,4), 448 blocks): u 1.15e-3, v 1.38e-3, and
    the rate from s=1 is P to within the reference's bounds (P - 0.3, P + 0.4)
    (sevenpatch_tests.jl:240-251)."""
    r1 = O.sevenpatch_errors(3, 2, roots=(2, 2))
    r = O.sevenpatch_errors(3, 2, roots=(4, 4))
    assert float(f"{r['l2u']:.3g}") == 1.15e-3 and float(f"{r['l2v']:.3g}") == 1.38e-3
    rate = T.convergence_rate([r1["h"], r["h"]], [r1["l2"], r["l2"]])
    assert 3 - 0.3 < rate < 3 + 0.4


def _tensor_truncate(v, P, E):
    """The flat tensor projector (I - topmode x topdual along each axis of every
    element) on a 3D per-block array v of shape (nb, N, N, N) (z, y, x)."""
    el = O.element(P)
    Pi = np.eye(P + 1) - np.outer(el["topmode"], el["topdual"])
    nb, N = v.shape[0], v.shape[1]
    w = v.reshape(nb, E, P + 1, E, P + 1, E, P + 1)
    w = np.einsum("ij,bzkyqxj->bzkyqxi", Pi, w)
    w = np.einsum("ij,bzkyjxl->bzkyixl", Pi, w)
    w = np.einsum("ij,bzjyqxl->bziyqxl", Pi, w)
    return w.reshape(nb, N, N, N)


@pytest.mark.parametrize("P", [2, 3])
def test_ball_truncation_is_metric_orthogonal_projection(P):
    """On the seven-patch ball the truncation is the sqrt(g)-orthogonal projection
    onto Q_{P-1} (sevenpatch_tests.jl:128-158): self-adjoint in the sqrt(g)-weighted
    quadrature, idempotent, its range left alone by the flat tensor projector
    (and a generic state not), u untouched, and no longer in that norm."""
    ref = O.Problem.ball(P, 2, roots=(2, 2))
    n, nn, N = ref.size(), ref.nn, ref.N
    rng = np.random.default_rng(31)
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    Pa, Pb = ref.truncate(a), ref.truncate(b)

    def nv(x):  # the sqrt(g)-weighted quadrature norm of v, normalized by the volume
        return ref.norms(x, vars=(2,))[0]

    def inner(x, y):
        return (nv(x + y) ** 2 - nv(x - y) ** 2) / 4

    scale = math.sqrt(inner(a, a) * inner(b, b))
    assert abs(inner(Pa, b) - inner(a, Pb)) < 1e-14 * scale
    assert np.allclose(ref.truncate(Pa), Pa, rtol=0, atol=1e-13 * np.max(np.abs(Pa)))
    va = Pa.reshape(ref.nb, 2, nn)[:, 1].reshape(ref.nb, N, N, N)
    assert np.max(np.abs(_tensor_truncate(va, P, 2) - va)) <= 1e-13 * np.max(np.abs(va))
    v0 = a.reshape(ref.nb, 2, nn)[:, 1].reshape(ref.nb, N, N, N)
    assert np.max(np.abs(_tensor_truncate(v0, P, 2) - v0)) > 1e-2
    assert np.array_equal(Pa.reshape(ref.nb, 2, nn)[:, 0], a.reshape(ref.nb, 2, nn)[:, 0])
    assert inner(Pa, Pa) <= inner(a, a)


# --- the annulus: a diagonal metric (geometry.jl:176-268, CODE.md:4172-4197) ------------
def test_annulus_published_table():
    """CODE.md:4188-4197 (the rotating Neumann mode m=2, n=1 on 1 <= r <= 2 to a
    quarter turn, roots (2s, s), E=2, cfl 1/2): P=3 truncated u 5.06e-3, 6.30e-4
    and v 3.35e-3, 4.23e-4 at s = 1, 2; P=4 truncated at s=1: u 3.77e-4, v 3.26e-4;
    P=2 without truncation at s=1: u 5.05e-3, v 3.77e-2; energy loss at most the
    table's 2.0e-10 (P=3).  And the wavenumber, k = 1.3406."""
    k = T.annulus_wavenumber(2, 1, 1.0, 2.0)
    assert float(f"{k:.5g}") == 1.3406
    rows = [(3, True, 1, 5.06e-3, 3.35e-3), (3, True, 2, 6.30e-4, 4.23e-4),
            (4, True, 1, 3.77e-4, 3.26e-4), (2, False, 1, 5.05e-3, 3.77e-2)]
    for P, trunc, s, eu, ev in rows:
        r = O.annulus_errors(P, 2, roots=(2 * s, s), m=2, k=k, truncate=trunc)
        assert float(f"{r['l2u']:.3g}") == eu and float(f"{r['l2v']:.3g}") == ev, (P, s, r)
        if P == 3 and s == 1:
            assert 0 < 1 - r["energy"] / r["energy0"] < 2.05e-10


@pytest.mark.parametrize("P", [2, 3])
def test_annulus_energy_rate_vanishes(P):
    """Every glue and wall of the annulus conserves the metric energy
    1/2 sum w (sqrt(g) v^2 + sum_d F_d^2 / A_d) of the semi-discrete system, with
    and without the truncation: the rate along f(y) is (E(y+f) - E(y-f))/2."""
    for trunc in (False, True):
        ref = O.Problem.annulus(P, 2, roots=(2, 2), truncate=trunc)
        y = ref.truncate(np.random.default_rng(7).standard_normal(ref.size()))
        f = ref.rhs(y)
        rate = (ref.energy(y + f) - ref.energy(y - f)) / 2
        assert abs(rate) < 1e-12 * (ref.energy(y) + ref.energy(f)), (P, trunc, rate)


@pytest.mark.parametrize("P", [2, 3, 4])
def test_annulus_truncation_is_metric_orthogonal_projection(P):
    """The per-line pairs make the truncation the sqrt(g)-orthogonal projection
    onto Q_{P-1} (evolution.jl:556-575): self-adjoint in the sqrt(g)-weighted
    quadrature, idempotent, u untouched, the flat tensor projector's range (and
    not its projection) kept, and no longer in that norm."""
    ref = O.Problem.annulus(P, 2, roots=(2, 2))
    n, nn = ref.size(), ref.nn
    rng = np.random.default_rng(32)
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    Pa, Pb = ref.truncate(a), ref.truncate(b)

    def nv(x):
        return ref.norms(x, vars=(2,))[0]

    def inner(x, y):
        return (nv(x + y) ** 2 - nv(x - y) ** 2) / 4

    scale = math.sqrt(inner(a, a) * inner(b, b))
    assert abs(inner(Pa, b) - inner(a, Pb)) < 1e-14 * scale
    assert np.allclose(ref.truncate(Pa), Pa, rtol=0, atol=1e-13 * np.max(np.abs(Pa)))
    assert np.array_equal(Pa.reshape(-1, 2, nn)[:, 0], a.reshape(-1, 2, nn)[:, 0])
    assert inner(Pa, Pa) <= inner(a, a)
    # the range is Q_{P-1} per element: the flat projector leaves it, not a generic state
    flat = O.Problem.brick(2, P, 2, roots=1)
    N = ref.N
    for y, keep in ((Pa, True), (a, False)):
        blk = y.reshape(-1, 2, nn)[:, 1]
        z = np.zeros((blk.shape[0], 2, nn))
        z[:, 1] = blk
        moved = max(np.max(np.abs(flat.truncate(zi.reshape(-1)) - zi.reshape(-1))) for zi in z)
        assert (moved < 1e-13 * np.max(np.abs(blk))) == keep, (keep, moved)
    # the angular lines' pair is the flat one (a constant factor), the radial one is not
    _, lm, ld = ref.dcoef()
    el = O.element(P)
    assert np.allclose(lm[:, 0], el["topmode"], rtol=0, atol=1e-15)
    assert np.allclose(ld[:, 0], el["topdual"], rtol=0, atol=1e-15)
    assert np.max(np.abs(lm[:, 1] - el["topmode"])) > 1e-6
