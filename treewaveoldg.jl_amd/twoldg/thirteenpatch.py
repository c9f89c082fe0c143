"""The thirteen-patch ball, its factored truncation, and the radiative boundary.

Mirrors TreeWaveOLDG.jl:
  ThirteenPatchBall, thirteenpatch_forest           geometry.jl:555-749
  factored_volume, radial_volume, angular_volume    geometry.jl:682-708
  FactoredTruncation / factored_truncation          evolution.jl:788-798, 904-984
  RadiativeBoundary, outer_radius                   evolution.jl:1129-1175
  OutgoingPulse, radiative_errors                   sevenpatch.jl:131-236

Patches (0-based): 0 the cube [-Lc, Lc]^3 of n roots per side; 1..6 the inner
shells over the cube's faces (k, s) in the seven-patch order, the seven-patch
ball's shells out to the middle sphere |x| = Rm, of m roots along k; 7..12 the
outer shells over the same faces, exactly radial out to |x| = R1, of m' roots
along k:

    c_k = s, c_j = 2 xi_j / n - 1 (j != k),  rho = xi_k/m' (s = 1), 1 - xi_k/m' (s = -1)
    x = (Rm + (R1 - Rm) rho) c / |c|

On an outer shell sqrt(g) = r^2 r' (2/n)^2 (1 + c_a^2 + c_b^2)^(-3/2), a radial
factor times an angular one that is the same function on all six, so the
sqrt(g)-orthogonal projection onto Q_{P-1} there is a rank-one pair along k
times a projection across it, whose angular tables are shared by every block
with the same tangential position and spacing.
"""
import math

import mpmath
import numpy as np

from . import _lib
from .annulus import _projector_pair
from .forest import Forest, refinement_regions
from .geometry import SevenPatchBall, shell_of, shell_index


class ThirteenPatchBall(SevenPatchBall):
    """The ball |x| <= R1 as a cube, six inner shells out to Rm and six exactly
    radial outer shells; roots = (n, m, m').  `factored = False` builds
    per-element tables on the outer shells too (the same projection, the
    reference tests' unfactored map)."""

    def __init__(self, Lc=1.0, Rm=2.0, R1=3.0, roots=(2, 1, 1), factored=True):
        if not Lc > 0:
            raise ValueError(f"the cube's half-width Lc must be positive, got {Lc}")
        if not Rm > math.sqrt(3) * Lc:
            raise ValueError(
                f"the middle radius Rm = {Rm} must exceed √3 Lc = {math.sqrt(3) * Lc}, so that "
                f"the cube's corners lie inside the middle sphere and every inner shell maps "
                f"one to one")
        if not R1 > Rm:
            raise ValueError(f"the outer radius R1 = {R1} must exceed the middle radius Rm = {Rm}")
        if len(roots) != 3:
            raise ValueError(
                f"the thirteen-patch ball takes roots = (n, m, m′), the cube's per side and the "
                f"inner and outer shells' radial ones, got {tuple(roots)!r}")
        if not all(r > 0 for r in roots):
            raise ValueError(
                f"every patch of the ball needs a positive number of roots along each "
                f"dimension, got {tuple(roots)}")
        self.Lc, self.Rm, self.R1 = float(Lc), float(Rm), float(R1)
        self.n, self.m, self.mo = (int(r) for r in roots)
        self.grading = 0.0
        self.factored = bool(factored)
        self._inner = SevenPatchBall(Lc, Rm, (self.n, self.m))

    def outer_radius(self):
        return self.R1

    def patch_roots(self, p):
        if p == 0:
            return (self.n,) * 3
        k, _ = shell_of(p if p <= 6 else p - 6)
        return tuple((self.m if p <= 6 else self.mo) if j == k else self.n for j in range(3))

    # -- the map ------------------------------------------------------------------
    def _outer(self, p, xi):
        k, s = shell_of(p - 6)
        c = np.empty_like(xi)
        for j in range(3):
            c[j] = float(s) if j == k else 2 * xi[j] / self.n - 1
        rho = xi[k] / self.mo if s == 1 else 1 - xi[k] / self.mo
        r = self.Rm + (self.R1 - self.Rm) * rho
        nc = np.sqrt(c[0] ** 2 + c[1] ** 2 + c[2] ** 2)
        return k, s, c, r, nc

    def position(self, p, xi):
        xi = np.asarray(xi, dtype=np.float64)
        if p <= 6:
            return self._inner.position(p, xi)
        _, _, c, r, nc = self._outer(p, xi)
        return (r / nc) * c

    def jacobian(self, p, xi):
        xi = np.asarray(xi, dtype=np.float64)
        if p <= 6:
            return self._inner.jacobian(p, xi)
        k, s, c, r, nc = self._outer(p, xi)
        M = np.zeros((xi.shape[1], 3, 3))
        dr = (self.R1 - self.Rm) * s / self.mo          # dr/dxi_k
        a = 2.0 / self.n                                 # dc_j/dxi_j
        for i in range(3):
            M[:, i, k] = dr * c[i] / nc
        for j in range(3):
            if j == k:
                continue
            for i in range(3):
                proj = (1.0 if i == j else 0.0) - c[i] * c[j] / nc ** 2
                M[:, i, j] = r * a / nc * proj
        return M

    # -- the factors of sqrt(g) on the outer shells (geometry.jl:682-708) ------------
    def factored_volume(self, p):
        """The radial dimension k (0-based) of patch p's factorization, or -1."""
        return shell_of(p - 6)[0] if p >= 7 else -1

    def radial_volume(self, xik, p):
        _, s = shell_of(p - 6)
        rho = xik / self.mo if s == 1 else 1 - xik / self.mo
        r = self.Rm + (self.R1 - self.Rm) * rho
        return r * r * ((self.R1 - self.Rm) / self.mo)

    def angular_volume(self, xa, xb):
        ca = 2 * xa / self.n - 1
        cb = 2 * xb / self.n - 1
        q = 1 + ca * ca + cb * cb
        return (2 / self.n) ** 2 / (q * np.sqrt(q))

    # -- glue ------------------------------------------------------------------------
    def neighbour_face(self, p, d, s):
        """The patch face glued to patch p's face (d, side s), or None on the outer
        sphere: cube to inner shells, inner to outer shells by the identity, and
        each layer's shells to each other by the quarter turns (geometry.jl:646-660)."""
        if p == 0:
            return shell_index(d, 1 if s else -1), d, 0 if s else 1
        outer = p >= 7
        q = p - 6 if outer else p
        k, sg = shell_of(q)
        up = (sg == 1 and s == 1) or (sg == -1 and s == 0)   # away from the cube
        if d == k:
            if outer:
                if up:
                    return None
                return q, k, 1 if sg == 1 else 0            # the inner shell's outer face
            if up:
                return q + 6, k, 0 if sg == 1 else 1        # the outer shell's inner face
            return 0, k, 1 if sg == 1 else 0
        t = 1 if s else -1
        return shell_index(d, t) + (6 if outer else 0), k, 1 if sg == 1 else 0

    # -- truncation tables --------------------------------------------------------------
    def _element_blocks(self, forest):
        return [b for b, lf in enumerate(forest.leaves)
                if lf[0] != 0 and not (self.factored and lf[0] >= 7)]

    def truncation_tables(self, forest, el):
        t = super().truncation_tables(forest, el)
        if self.factored:
            t.update(self.factored_tables(forest, el))
        return t

    def factored_tables(self, forest, el):
        """The factored truncation of the outer shells (evolution.jl:904-984):
        fslot (nb; -1 off the outer shells), fradial (nf), the radial pairs
        frmodes/frduals (nf, E, P+1) along k, faslot (nf) into the angular tables
        famodes ((P+1)^2, P^2) and, per table and element across k, faweights
        (nA, E^2, (P+1)^2) and the upper Cholesky factor fachol (nA, E^2, P^4) of
        G = Psi^T W Psi, column-major."""
        P, P1, E, N = el.P, el.P1, forest.E, forest.N
        h = forest.spacings()
        blocks = [b for b, lf in enumerate(forest.leaves) if lf[0] >= 7]
        nb = forest.nblocks
        fslot = np.full(nb, -1, dtype=np.int32)
        fslot[blocks] = np.arange(len(blocks), dtype=np.int32)
        nf = len(blocks)
        fradial = np.array([self.factored_volume(forest.leaves[b][0]) for b in blocks],
                           dtype=np.int32)
        rmodes = np.empty((nf, E, P1))
        rduals = np.empty((nf, E, P1))
        keys = {}
        aslot = np.empty(nf, dtype=np.int32)
        first = []
        for f, b in enumerate(blocks):
            p, lev, *c = forest.leaves[b]
            k = int(fradial[f])
            across = tuple(c[j] * N * h[b] for j in range(3) if j != k)
            key = (across, h[b])
            if key not in keys:
                keys[key] = len(keys)
                first.append(b)
            aslot[f] = keys[key]
        with mpmath.workdps(40):
            L = [[mpmath.legendre(q, mpmath.mpf(float(xi))) for q in range(P1)]
                 for xi in el.nodes]
            w = [mpmath.mpf(float(x)) for x in el.weights]
            off = [mpmath.mpf(float(x)) for x in el.offsets]
            cache = {}
            for f, b in enumerate(blocks):
                p, lev, *c = forest.leaves[b]
                k = int(fradial[f])
                lo = forest.patches[p][1][k]
                for e in range(E):
                    xs = [mpmath.mpf(lo) + mpmath.mpf(float(h[b])) *
                          (mpmath.mpf(c[k] * N + e * P1) + off[i]) for i in range(P1)]
                    a = tuple(self._radial_volume_mp(x, p) for x in xs)
                    if a not in cache:
                        cache[a] = _projector_pair(L, w, a, P)
                    rmodes[f, e], rduals[f, e] = cache[a]
        # angular tables: Psi over the (P+1)^2 nodes of a plane, P^2 modes
        Lg = el.legendre
        k0, k1 = np.meshgrid(np.arange(P1), np.arange(P1), indexing="ij")
        q0, q1 = np.meshgrid(np.arange(P), np.arange(P), indexing="ij")
        kk = [a.T.reshape(-1) for a in (k0, k1)]        # node n = k0 + P1 k1
        qq = [a.T.reshape(-1) for a in (q0, q1)]        # mode q = q0 + P q1
        Psi = Lg[kk[0][:, None], qq[0][None, :]] * Lg[kk[1][:, None], qq[1][None, :]]
        wn = el.weights[kk[0]] * el.weights[kk[1]]
        nA, m2 = len(first), P * P
        aweights = np.empty((nA, E * E, P1 * P1))
        achol = np.empty((nA, E * E, m2 * m2))
        for A, b in enumerate(first):
            p, lev, *c = forest.leaves[b]
            k = self.factored_volume(p)
            tj = [j for j in range(3) if j != k]
            lo = forest.patches[p][1]
            for ea in range(E * E):
                e = (ea % E, ea // E)
                xa = lo[tj[0]] + h[b] * (c[tj[0]] * N + e[0] * P1 + el.offsets[kk[0]])
                xb = lo[tj[1]] + h[b] * (c[tj[1]] * N + e[1] * P1 + el.offsets[kk[1]])
                wts = wn * self.angular_volume(xa, xb)
                aweights[A, ea] = wts
                G = Psi.T @ (wts[:, None] * Psi)
                Lc = np.linalg.cholesky(G)                  # G = L L^T, U = L^T
                achol[A, ea] = Lc.reshape(-1)               # U column-major = L row-major
        return dict(fslot=fslot, nfactored=nf, fradial=fradial,
                    frmodes=np.ascontiguousarray(rmodes), frduals=np.ascontiguousarray(rduals),
                    faslot=aslot, nangular=nA, famodes=np.ascontiguousarray(Psi),
                    faweights=np.ascontiguousarray(aweights),
                    fachol=np.ascontiguousarray(achol))

    def _radial_volume_mp(self, xik, p):
        _, s = shell_of(p - 6)
        mo = mpmath.mpf(self.mo)
        rho = xik / mo if s == 1 else 1 - xik / mo
        R1, Rm = mpmath.mpf(self.R1), mpmath.mpf(self.Rm)
        r = Rm + (R1 - Rm) * rho
        return r * r * ((R1 - Rm) / mo)


def thirteenpatch_forest(P, E, roots=(2, 1, 1), Lc=1.0, Rm=2.0, R1=3.0, refined=False,
                         factored=True):
    """The thirteen-patch ball as a forest and its map (geometry.jl:712-749)."""
    N = E * (P + 1)
    if E % 2:
        raise ValueError(
            f"E = {E} elements of order P = {P} per block is odd, and E must be even: TreeAMR's "
            f"element centering needs N = {N} to be a multiple of 2(P+1) = {2 * (P + 1)}")
    geo = ThirteenPatchBall(Lc, Rm, R1, tuple(roots), factored=factored)
    patches = []
    for p in range(13):
        r = geo.patch_roots(p)
        patches.append((r, (0.0,) * 3, tuple(float(x) for x in r)))
    forest = Forest(3, P, E, patches, periodic=(False,) * 3, reflecting=False, geometry=geo)
    if refined is not False:
        forest.refine_regions(refinement_regions(refined, 3, geo.n), geo.n, patch=0)
        for lf in forest.leaves:
            p, lev, *c = lf
            on_face = p == 0 and lev > 0 and any(
                c[d] == 0 or c[d] == (geo.n << lev) - 1 for d in range(3))
            if on_face or (p != 0 and lev != 0):
                raise ValueError(
                    f"the refinement {refined!r} reaches the cube's faces, and the glue keeps "
                    f"the leaves across a glued face at one level, so it would refine the "
                    f"shell across it, whose map is not affine: keep the regions at least one "
                    f"root inside the cube [0, {geo.n}]^3")
    return forest, geo


class RadiativeBoundary:
    """n . grad u = -(kappa du/dt + u/R) on a full geometry's outer faces, R the
    geometry's outer radius; inverse_r = False drops u/R (evolution.jl:1129-1164).
    Imposed through the face's ghosts: u's ghost is the face node's own u, the
    gradient's makes the central flux -sigma |S| (kappa v + u/R)."""

    def __init__(self, kappa=1.0, inverse_r=True):
        if not kappa >= 0:
            raise ValueError(
                f"κ = {kappa} is negative, and a radiative boundary with κ < 0 feeds energy in: "
                f"take κ ≥ 0, κ = 1 for Sommerfeld's condition")
        self.kappa = float(kappa)
        self.inverse_r = bool(inverse_r)

    def invR(self, geometry):
        if not self.inverse_r:
            return 0.0
        R = getattr(geometry, "outer_radius", None)
        if R is None:
            raise ValueError(
                f"the geometry {type(geometry).__name__} defines no outer radius, which the u/R "
                f"term of a radiative boundary needs: define `outer_radius` for it, or pass "
                f"`RadiativeBoundary(inverse_r=False)`")
        return 1.0 / R()


class OutgoingPulse:
    """u = (f(t - r) - f(t + r))/r, f(s) = exp(-s^2/w^2): regular at the origin,
    outgoing up to exp(-R^2/w^2) on a sphere of radius R (sevenpatch.jl:131-176).
    Below r = w/1000 the quotient is replaced by its Taylor series through r^4."""

    def __init__(self, w):
        self.w = float(w)

    def __call__(self, x, t):
        w = self.w
        r = np.sqrt(np.sum(np.asarray(x) ** 2, axis=0))

        def f(s):
            return np.exp(-(s / w) ** 2)

        def df(s):
            return -2 * s / w ** 2 * f(s)

        far = r > w / 1000
        rr = np.where(far, r, 1.0)
        u = np.where(far, (f(t - rr) - f(t + rr)) / rr, 0.0)
        v = np.where(far, (df(t - rr) - df(t + rr)) / rr, 0.0)
        if not np.all(far):
            y = t / w
            H = (2 * y, 4 * y ** 2 - 2, 8 * y ** 3 - 12 * y, 16 * y ** 4 - 48 * y ** 2 + 12,
                 32 * y ** 5 - 160 * y ** 3 + 120 * y,
                 64 * y ** 6 - 480 * y ** 4 + 720 * y ** 2 - 120)

            def dd(n):
                return (-1 / w) ** n * H[n - 1] * math.exp(-(t / w) ** 2)

            r2 = r * r
            us = -2 * (dd(1) + r2 * dd(3) / 6 + r2 ** 2 * dd(5) / 120)
            vs = -2 * (dd(2) + r2 * dd(4) / 6 + r2 ** 2 * dd(6) / 120)
            u = np.where(far, u, us)
            v = np.where(far, v, vs)
        return u, v


def pulse_state(pulse, t):
    """(x, var) -> value of the pulse at physical x (PulseState, sevenpatch.jl:178-185)."""
    def f(x, var):
        return pulse(x, t)[var - 1]
    return f


def ZeroData():
    """Zero Dirichlet data on the outer sphere: the plane wave of k = 0, phi = 0."""
    from .drivers import PlaneWave
    return PlaneWave((0.0, 0.0, 0.0), 0.0)
