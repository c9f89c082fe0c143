"""The four-patch annulus: a diagonal metric (TreeWaveOLDG.jl src/geometry.jl,
src/annulus.jl, src/evolution.jl).

  PolarAnnulus (sqrt(g), A_1, A_2, jacobian_factor)   geometry.jl:176-238
  annulus_forest (four patches glued in theta)        geometry.jl:240-268
  metric_truncation / projector_mode                   evolution.jl:576-622
  block_speeds (hspeed)                                evolution.jl:643-656, 698-708
  annulus_wavenumber, AnnulusMode, annulus_exact      annulus.jl:14-135
  annulus_errors                                       annulus.jl:137-162

Patch p (0-based) maps its logical box [0, n1] x [0, n2] to
theta = p pi/2 + theta' x_1, r = r0 + r' x_2, theta' = (pi/2)/n1,
r' = (r1 - r0)/n2, so sqrt(g) = r theta' r', A_1 = r'/(r theta'),
A_2 = r theta'/r'.  Each patch's high angular face is glued by the identity to
the next patch's low one; the radial faces are reflecting walls.  The
right-hand side's gradient sets hold the fluxes F_d = A_d D_d u and the
Laplacian divides by sqrt(g); the truncation is the tensor product of per-line
rank-one projectors orthogonal in the weights times the factor of sqrt(g)
along the line, which is the sqrt(g)-orthogonal projection because sqrt(g)
factorizes.  All of it runs in libtwoldg; this module builds the tables on
the host and evaluates the exact mode (Bessel functions, scipy).
"""
import math

import mpmath
import numpy as np

from . import _lib
from .forest import Forest


class PolarAnnulus:
    """The annulus r0 <= r <= r1 as four patches of `roots` = (n1, n2) roots;
    the map's constants theta' and r' built at 512 bits and rounded once."""

    kind = "diagonal"
    outer_faces = False  # every outer face is a wall: no boundary data

    def __init__(self, r0, r1, roots):
        if not 0 < r0 < r1:
            raise ValueError(f"an annulus needs 0 < r0 < r1, got r0 = {r0} and r1 = {r1}")
        if not all(int(n) > 0 for n in roots):
            raise ValueError(
                f"every patch of the annulus needs a positive number of roots along each "
                f"dimension, got {tuple(roots)}")
        self.r0, self.r1 = float(r0), float(r1)
        self.roots = (int(roots[0]), int(roots[1]))
        with mpmath.workprec(512):
            self.dth = float(mpmath.pi / 2 / self.roots[0])
            self.dr = float((mpmath.mpf(r1) - mpmath.mpf(r0)) / self.roots[1])
            self.halfpi = float(mpmath.pi / 2)

    # -- the map -------------------------------------------------------------
    def radius(self, x2):
        return self.r0 + self.dr * x2

    def angle(self, x1, p):
        return p * self.halfpi + self.dth * x1

    def physical(self, forest, x):
        """Physical positions (nb, 2, N^2) of logical positions x (nb, 2, N^2)."""
        out = np.empty_like(x)
        for b, (p, lev, *c) in enumerate(forest.leaves):
            r = self.radius(x[b, 1])
            th = self.angle(x[b, 0], p)
            out[b, 0] = r * np.cos(th)
            out[b, 1] = r * np.sin(th)
        return out

    # -- glue ------------------------------------------------------------------
    def glue_face(self, forest, b, d, s, F):
        """A patch-boundary face: the angular faces to the next (previous) patch,
        same level, by the identity; the radial faces walls."""
        p, lev, *c = forest.leaves[b]
        if d == 1:
            F.kind = _lib.FACE_MIRROR
            return
        q = (p + (1 if s else 3)) % 4
        cc = list(c)
        cc[0] = 0 if s else forest.dims(q, lev)[0] - 1
        key = (q, lev) + tuple(cc)
        if key not in forest.index:
            raise ValueError("the annulus takes a uniform forest (no refinement across glues)")
        F.kind = _lib.FACE_SAME
        F.nb[0] = forest.index[key]

    # -- what a diagonal metric adds to a problem ---------------------------------
    def metric(self, forest, el):
        """(sqrt(g), A_1, A_2) at the nodes, (nb, 3, N^2), and hspeed = min over
        blocks of h_b over the block's largest logical speed max_d sqrt(A_d/sqrt(g))."""
        x = forest.node_logical(el.offsets)
        r = self.radius(x[:, 1])
        J = r * self.dth * self.dr
        A1 = self.dr / (r * self.dth)
        A2 = r * self.dth / self.dr
        coef = np.ascontiguousarray(np.stack([J, A1, A2], axis=1))
        speed = np.maximum(np.sqrt(A1 / J), np.sqrt(A2 / J)).max(axis=1)
        hspeed = float(np.min(forest.spacings() / speed))
        return dict(dcoef=coef, hspeed=hspeed)

    def jacobian_factor(self, d, x):
        """The factor of sqrt(g) along d at logical coordinate x (an mpf)."""
        if d == 0:
            return mpmath.mpf(self.dth)
        return (mpmath.mpf(self.r0) + mpmath.mpf(self.dr) * x) * mpmath.mpf(self.dr)

    def truncation_tables(self, forest, el):
        """Per block, dimension and element along it the rank-one pair (mode,
        dual) of the line's projector, (nb, D, E, P+1) each (evolution.jl:596-622)."""
        P, P1, E, N = el.P, el.P1, forest.E, forest.N
        nb = forest.nblocks
        mode = np.empty((nb, 2, E, P1))
        dual = np.empty((nb, 2, E, P1))
        h = forest.spacings()
        cache = {}
        with mpmath.workdps(40):
            L = [[mpmath.legendre(k, mpmath.mpf(float(xi))) for k in range(P1)] for xi in el.nodes]
            w = [mpmath.mpf(float(x)) for x in el.weights]
            off = [mpmath.mpf(float(x)) for x in el.offsets]
            for b, (p, lev, *c) in enumerate(forest.leaves):
                lo = forest.patches[p][1]
                for d in range(2):
                    for e in range(E):
                        a = tuple(self.jacobian_factor(
                            d, mpmath.mpf(lo[d]) + mpmath.mpf(float(h[b])) *
                            (mpmath.mpf(c[d] * N + e * P1) + off[i])) for i in range(P1))
                        if a not in cache:
                            cache[a] = _projector_pair(L, w, a, P)
                        mode[b, d, e], dual[b, d, e] = cache[a]
        return dict(lmode=np.ascontiguousarray(mode), ldual=np.ascontiguousarray(dual))


def _projector_pair(L, w, a, P):
    """The degree-P polynomial at the nodes orthogonal to every lower one in
    sum w_i a_i f_i g_i -- L_P less its projection on L_0 .. L_{P-1} -- and the
    dual w a phi / gamma, gamma = sum w a phi^2 (evolution.jl:576-589)."""
    P1 = P + 1

    def inner(k, l):
        return mpmath.fsum(w[i] * a[i] * L[i][k] * L[i][l] for i in range(P1))

    G = mpmath.matrix([[inner(k, l) for l in range(P)] for k in range(P)])
    rhs = mpmath.matrix([inner(P, k) for k in range(P)])
    cvec = mpmath.lu_solve(G, rhs)
    phi = [L[i][P] - mpmath.fsum(cvec[k] * L[i][k] for k in range(P)) for i in range(P1)]
    gam = mpmath.fsum(w[i] * a[i] * phi[i] ** 2 for i in range(P1))
    return (np.array([float(x) for x in phi]),
            np.array([float(w[i] * a[i] * phi[i] / gam) for i in range(P1)]))


def annulus_forest(P, E, roots=(2, 1), r0=1.0, r1=2.0):
    """The annulus as four patches of `roots` = (n1, n2) roots each holding E^2
    elements of order P per block, and its PolarAnnulus map (geometry.jl:240-268)."""
    N = E * (P + 1)
    if E % 2:
        raise ValueError(
            f"E = {E} elements of order P = {P} per block is odd, and E must be even: "
            f"TreeAMR's element centering needs N = {N} to be a multiple of 2(P+1) = "
            f"{2 * (P + 1)}")
    geo = PolarAnnulus(r0, r1, roots)
    n1, n2 = geo.roots
    patches = [((n1, n2), (0.0, 0.0), (float(n1), float(n2))) for _ in range(4)]
    forest = Forest(2, P, E, patches, periodic=(False, False), reflecting=True, geometry=geo)
    return forest, geo


# -- the rotating Bessel mode (annulus.jl) ---------------------------------------------
def _bessel():
    from scipy.special import jv, yv
    return jv, yv


def _dj(m, z):
    jv, _ = _bessel()
    return (jv(m - 1, z) - jv(m + 1, z)) / 2


def _dy(m, z):
    _, yv = _bessel()
    return (yv(m - 1, z) - yv(m + 1, z)) / 2


def annulus_cross(m, k, r0, r1):
    """J_m'(k r0) Y_m'(k r1) - J_m'(k r1) Y_m'(k r0) (annulus.jl:20-21)."""
    return _dj(m, k * r0) * _dy(m, k * r1) - _dj(m, k * r1) * _dy(m, k * r0)


def annulus_wavenumber(m, n, r0, r1):
    """The n-th positive root k of annulus_cross: the wavenumber (and frequency)
    of the Neumann mode of angular number m and radial number n on r0 <= r <= r1.
    Bracketed by a scan in steps of a fiftieth of pi/(r1 - r0) from an eighth of
    a step, and bisected to the last bit (annulus.jl:23-65)."""
    if not (m >= 0 and n >= 1):
        raise ValueError(f"the mode needs m >= 0 and n >= 1, got m = {m} and n = {n}")
    r0, r1 = float(r0), float(r1)
    if not 0 < r0 < r1:
        raise ValueError("the annulus needs 0 < r0 < r1")

    def f(k):
        return float(annulus_cross(m, k, r0, r1))

    step = math.pi / (r1 - r0) / 50
    a = step / 8
    fa = f(a)
    found = 0
    while True:
        b = a + step
        fb = f(b)
        if fa == 0 or math.copysign(1, fa) != math.copysign(1, fb):
            found += 1
            if found == n:
                lo, hi, flo = a, b, fa
                while True:
                    mid = (lo + hi) / 2
                    if mid == lo or mid == hi:
                        return lo
                    fm = f(mid)
                    if fm == 0:
                        return mid
                    if math.copysign(1, fm) == math.copysign(1, flo):
                        lo, flo = mid, fm
                    else:
                        hi = mid
        a, fa = b, fb


def annulus_exact(geo, forest, m, k, t):
    """u = R(kr) cos(m theta - k t), v = k R(kr) sin(m theta - k t) with
    R(z) = s (J_m(z) Y_m'(k r0) - Y_m(z) J_m'(k r0)), s = pi k r0 / 2 so that
    R(k r0) = 1, at the logical positions of each patch (annulus.jl:67-135):
    f(x, var) for fill_at_nodes, x (2, n) the logical positions of the blocks in
    forest order."""
    jv, yv = _bessel()
    k = float(k)
    jr0, yr0 = _dj(m, k * geo.r0), _dy(m, k * geo.r0)
    scale = math.pi * k * geo.r0 / 2
    N2 = forest.N ** 2
    patch = np.repeat(np.array([lf[0] for lf in forest.leaves]), N2)

    def f(x, var):
        r = geo.radius(x[1])
        th = geo.angle(x[0], patch)
        R = scale * (jv(m, k * r) * yr0 - yv(m, k * r) * jr0)
        ph = m * th - k * t
        return R * np.cos(ph) if var == 1 else k * R * np.sin(ph)
    return f


def annulus_errors(P, E, roots=(2, 1), r0=1.0, r1=2.0, m=1, n=1, cfl=0.5, tend=None,
                   truncate=True, precision="f64", nsteps=None):
    """Evolve the rotating Neumann mode on the four-patch annulus to tend (by
    default a quarter turn of its pattern, (pi/2) m / k) and return the
    sqrt(g)-weighted L2 errors of u and v together and apart, the nodal L-inf
    error, the physical radial spacing h and element width, the step count,
    the blocks, the wavenumber and the energies (annulus.jl:137-162)."""
    from .element import Element
    from .problem import (WaveProblem, initial_state, fill_at_nodes_forest, oldg_timestep,
                          wave_integrator, wave_energy)
    from .drivers import _errors
    forest, geo = annulus_forest(P, E, roots=roots, r0=r0, r1=r1)
    el = Element(P)
    problem = WaveProblem(forest, el, truncate=truncate, precision=precision, geometry=geo)
    k = annulus_wavenumber(m, n, r0, r1)
    t_end = math.pi / 2 * max(m, 1) / k if tend is None else float(tend)
    u0 = initial_state(problem, annulus_exact(geo, forest, m, k, 0.0))
    if nsteps is None:
        nsteps = math.ceil(t_end / oldg_timestep(problem, cfl))
    dt = t_end / nsteps
    energy0 = wave_energy(problem, u0)
    integ = wave_integrator(problem, u0, 0.0, t_end, dt)
    integ.solve()
    energy = integ.energy()
    uexact = fill_at_nodes_forest(forest, annulus_exact(geo, forest, m, k, t_end), el)
    problem.upload(uexact, _lib.SLOT_AUX)
    l2, linf = _errors(problem, (1, 2))
    l2u, _ = _errors(problem, (1,))
    l2v, _ = _errors(problem, (2,))
    h = forest.minimum_spacing() * geo.dr
    return dict(l2=l2, l2u=l2u, l2v=l2v, linf=linf, h=h, he=(P + 1) * h, nsteps=nsteps, dt=dt,
                tend=t_end, nblocks=forest.nblocks, k=k, energy0=energy0, energy=energy,
                problem=problem, integrator=integ)
