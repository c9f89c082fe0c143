"""ctypes front end of the OLDG oracle (test infrastructure only).

The oracle is a CPU FP64 restatement of TreeWaveOLDG.jl's field-evolution path
(see oldg_oracle.cpp for the file:line of each part it follows).  Only the
tests, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg use it,
and only as the checker; the product never imports it.

Drivers mirror the reference's: ``wave_errors`` follows sinewave.jl:228-268,
``sevenpatch_errors`` sevenpatch.jl:84-117.
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.c_void_p


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboldg_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.orc_element.argtypes = [ctypes.c_int, _dp]
        L.orc_brick.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                                      ctypes.c_int, _dp]
        L.orc_brick.restype = _dp
        L.orc_ball.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] * 3 + [ctypes.c_int, _dp,
                                                                             ctypes.c_double,
                                                                             ctypes.c_int, _dp]
        L.orc_ball.restype = _dp
        L.orc_annulus.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] * 2 + [ctypes.c_int]
        L.orc_annulus.restype = _dp
        L.orc_dcoef.argtypes = [_dp, _dp, _dp, _dp]
        L.orc_free.argtypes = [_dp]
        L.orc_nblocks.argtypes = [_dp]
        L.orc_nblocks.restype = ctypes.c_int64
        L.orc_N.argtypes = [_dp]
        L.orc_hspeed.argtypes = [_dp]
        L.orc_hspeed.restype = ctypes.c_double
        L.orc_block.argtypes = [_dp, ctypes.c_int64, _dp, _dp]
        L.orc_faces.argtypes = [_dp, _dp, _dp]
        L.orc_coords.argtypes = [_dp, _dp]
        L.orc_coef.argtypes = [_dp, _dp, _dp]
        L.orc_fill_exact.argtypes = [_dp, ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.orc_rhs.argtypes = [_dp, _dp, _dp, ctypes.c_double]
        L.orc_truncate.argtypes = [_dp, _dp]
        L.orc_energy.argtypes = [_dp, _dp, ctypes.c_double]
        L.orc_energy.restype = ctypes.c_double
        L.orc_gradients.argtypes = [_dp, _dp, ctypes.c_double, _dp]
        L.orc_ghosts.argtypes = [_dp, _dp, ctypes.c_double, _dp, _dp]
        L.orc_divergence.argtypes = [_dp, _dp, ctypes.c_int, _dp]
        L.orc_norms.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_min_spacing.argtypes = [_dp]
        L.orc_min_spacing.restype = ctypes.c_double
        L.orc_timestep.argtypes = [_dp, ctypes.c_double]
        L.orc_timestep.restype = ctypes.c_double
        L.orc_rk4.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int64]
        L.orc_threads.restype = ctypes.c_int
        L.orc_ball_ext.argtypes = [ctypes.c_int] * 5 + [ctypes.c_double] * 3 + [
            ctypes.c_int, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_ball_ext.restype = _dp
        L.orc_brick_leaves.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double, ctypes.c_int,
                                                             ctypes.c_int, ctypes.c_int64, _dp]
        L.orc_brick_leaves.restype = _dp
        L.orc_modes.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                ctypes.c_double, _dp]
        L.orc_transfer.argtypes = [_dp, _dp, _dp, _dp, _dp]
        L.orc_duals.argtypes = [ctypes.c_int, _dp]
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def element(P):
    """Tables in binary128 rounded once: dict of nodes, weights, offsets,
    stencil, trunc, topmode, topdual (element.jl:129-170)."""
    out = np.zeros(600)
    lib().orc_element(P, _ptr(out))
    P1 = P + 1
    k = 0
    res = {}
    for name, shape in [("nodes", (P1,)), ("weights", (P1,)), ("offsets", (P1,)),
                        ("stencil", (P1, P1 + 2)), ("trunc", (P1, P1)), ("topmode", (P1,)),
                        ("topdual", (P1,))]:
        n = int(np.prod(shape))
        res[name] = out[k:k + n].reshape(shape).copy()
        k += n
    return res


def duals(P):
    """The Legendre duals (P+1, P+1) [mode][node] in binary128 rounded once."""
    out = np.zeros((P + 1) * (P + 1))
    lib().orc_duals(P, _ptr(out))
    return out.reshape(P + 1, P + 1)


class Problem:
    """A forest plus everything the RHS needs (the oracle's WaveProblem)."""

    def __init__(self, handle, D, P, E, geometry):
        self.h = handle
        self.D, self.P, self.E = D, P, E
        self.N = E * (P + 1)
        self.nn = self.N ** D
        self.nb = lib().orc_nblocks(handle)
        self.geometry = geometry

    @classmethod
    def brick(cls, D, P, E, roots=4, L=1.0, reflecting=False, truncate=True, regions=()):
        reg = np.asarray(regions, dtype=np.float64).reshape(-1)
        nreg = len(regions)
        h = lib().orc_brick(D, P, E, roots, float(L), int(reflecting), int(truncate), nreg,
                            _ptr(reg) if nreg else None)
        return cls(h, D, P, E, 0)

    @classmethod
    def ball(cls, P, E, roots=(2, 2), Lc=1.0, R1=3.0, grading=0.0, truncate=True,
             k=(1.1, -0.7, 0.5), phi=0.3, regions=()):
        kv = np.asarray(k, dtype=np.float64)
        reg = np.asarray(regions, dtype=np.float64).reshape(-1)
        h = lib().orc_ball(P, E, roots[0], roots[1], float(Lc), float(R1), float(grading),
                           int(truncate), _ptr(kv), float(phi), len(regions),
                           _ptr(reg) if len(regions) else None)
        p = cls(h, 3, P, E, 1)
        p.k = kv
        p.phi = phi
        return p

    @classmethod
    def ball_ext(cls, P, E, roots=(2, 2), Lc=1.0, Rm=2.0, R1=3.0, truncate=True,
                 k=(1.1, -0.7, 0.5), phi=0.3, radiative=False, kappa=1.0, invR=0.0):
        """The seven-patch ball (roots (n, m)) or the thirteen-patch ball (roots
        (n, m, m'), inner shells to Rm), with a plane wave's Dirichlet data or the
        radiative boundary on the outer sphere."""
        kv = np.asarray(k, dtype=np.float64)
        mo = roots[2] if len(roots) == 3 else 0
        h = lib().orc_ball_ext(P, E, roots[0], roots[1], mo, float(Lc), float(Rm), float(R1),
                               int(truncate), _ptr(kv), float(phi), int(radiative),
                               float(kappa), float(invR))
        p = cls(h, 3, P, E, 1)
        p.k = kv
        p.phi = phi
        return p

    @classmethod
    def annulus(cls, P, E, roots=(2, 1), r0=1.0, r1=2.0, truncate=True):
        """The four-patch annulus with its diagonal metric (geometry.jl:176-268)."""
        h = lib().orc_annulus(P, E, roots[0], roots[1], float(r0), float(r1), int(truncate))
        return cls(h, 2, P, E, 2)

    @classmethod
    def brick_leaves(cls, D, P, E, leaves, roots=4, L=1.0, reflecting=False, truncate=True):
        """A brick forest of explicit leaves, (level, corner) pairs in any order."""
        a = np.zeros((len(leaves), 4), dtype=np.int64)
        for i, (lev, c) in enumerate(leaves):
            a[i, 0] = lev
            a[i, 1:1 + len(c)] = c
        h = lib().orc_brick_leaves(D, P, E, roots, float(L), int(reflecting), int(truncate),
                                   len(leaves), _ptr(a))
        return cls(h, D, P, E, 0)

    def modes(self, y, var, modes, weights, pows, post):
        """Per element post * max over lines and q of ((weights[q]/A) |c_modes[q]|)^pows[q]
        of variable var (0 = u, 1 = v), shape (nb, E^D) with e_1 fastest
        (refinement.jl:56-101)."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        m = np.asarray(modes, dtype=np.int32)
        w = np.asarray(weights, dtype=np.float64)
        pw = np.asarray(pows, dtype=np.float64)
        tau = np.zeros((self.nb, self.E ** self.D))
        lib().orc_modes(self.h, _ptr(y), int(var), len(m), _ptr(m), _ptr(w), _ptr(pw),
                        float(post), _ptr(tau))
        return tau

    def transfer_from(self, src, ysrc):
        """This forest's state from src's ysrc by keys (copy, refine, coarsen); returns
        the state and the number of leaves of each kind."""
        ysrc = np.ascontiguousarray(ysrc, dtype=np.float64)
        y = np.zeros(self.size())
        counts = np.zeros(3, dtype=np.int64)
        if lib().orc_transfer(self.h, _ptr(y), src.h, _ptr(ysrc), _ptr(counts)):
            raise ValueError("a leaf of the regridded forest has no origin in the source")
        return y, tuple(int(c) for c in counts)

    def dcoef(self):
        """(sqrt(g), A_1, A_2) at the nodes (nb, 3, nn) and the truncation pairs
        (nb, D, E, P+1) of an annulus."""
        c = np.zeros(self.nb * 3 * self.nn)
        m = np.zeros(self.nb * 2 * self.E * (self.P + 1))
        d = np.zeros_like(m)
        if not lib().orc_dcoef(self.h, _ptr(c), _ptr(m), _ptr(d)):
            return None
        sh = (self.nb, 2, self.E, self.P + 1)
        return c.reshape(self.nb, 3, self.nn), m.reshape(sh), d.reshape(sh)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    def size(self):
        return self.nb * 2 * self.nn

    def blocks(self):
        info = np.zeros(5, dtype=np.int64)
        hh = np.zeros(1)
        out = []
        for b in range(self.nb):
            lib().orc_block(self.h, b, _ptr(info), _ptr(hh))
            out.append((int(info[0]), int(info[1]), tuple(int(x) for x in info[2:2 + self.D]),
                        float(hh[0])))
        return out

    def faces(self):
        kind = np.zeros(self.nb * 2 * self.D, dtype=np.int64)
        nb = np.zeros(self.nb * 2 * self.D, dtype=np.int64)
        lib().orc_faces(self.h, _ptr(kind), _ptr(nb))
        return kind.reshape(self.nb, self.D, 2), nb.reshape(self.nb, self.D, 2)

    def coords(self):
        x = np.zeros(self.nb * self.D * self.nn)
        lib().orc_coords(self.h, _ptr(x))
        return x.reshape(self.nb, self.D, self.nn)

    def coef(self):
        c = np.zeros(self.nb * (1 + self.D * self.D) * self.nn)
        f = np.zeros(self.nb, dtype=np.uint8)
        if not lib().orc_coef(self.h, _ptr(c), _ptr(f)):
            return None, None
        return c.reshape(self.nb, 1 + self.D * self.D, self.nn), f.astype(bool)

    def fill_exact(self, kind, t, prm=()):
        y = np.zeros(self.size())
        p = np.asarray(prm, dtype=np.float64)
        lib().orc_fill_exact(self.h, kind, float(t), _ptr(p) if len(p) else None, _ptr(y))
        return y

    def rhs(self, y, t=0.0):
        y = np.ascontiguousarray(y, dtype=np.float64)
        du = np.zeros_like(y)
        lib().orc_rhs(self.h, _ptr(y), _ptr(du), float(t))
        return du

    def truncate(self, y):
        y = np.array(y, dtype=np.float64, copy=True)
        lib().orc_truncate(self.h, _ptr(y))
        return y

    def energy(self, y, t=0.0):
        y = np.ascontiguousarray(y, dtype=np.float64)
        return lib().orc_energy(self.h, _ptr(y), float(t))

    def gradients(self, y, t=0.0):
        y = np.ascontiguousarray(y, dtype=np.float64)
        g = np.zeros(self.D * self.nb * self.nn)
        lib().orc_gradients(self.h, _ptr(y), float(t), _ptr(g))
        return g

    def ghosts(self, y, t=0.0):
        """(u ghosts, gradient ghosts) in rows (b, d, s, t) after fill_gradient."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        nt = self.N ** (self.D - 1)
        rows = self.nb * 2 * self.D * nt
        ug = np.zeros(rows)
        gg = np.zeros(rows * (self.D if self.geometry else 1))
        lib().orc_ghosts(self.h, _ptr(y), float(t), _ptr(ug), _ptr(gg))
        return ug, gg

    def divergence(self, g, d):
        """D_d g with g exchanged as the gradient's component d (odd under a
        mirror in d); g and the result are (nblocks, N^D)."""
        g = np.ascontiguousarray(g, dtype=np.float64)
        out = np.zeros(self.nb * self.nn)
        lib().orc_divergence(self.h, _ptr(g), int(d), _ptr(out))
        return out.reshape(self.nb, self.nn)

    def norms(self, err, vars=(1, 2), p=2):
        err = np.ascontiguousarray(err, dtype=np.float64)
        mask = sum(1 << (v - 1) for v in vars)
        lp = np.zeros(1)
        li = np.zeros(1)
        lib().orc_norms(self.h, _ptr(err), mask, int(p), _ptr(lp), _ptr(li))
        return float(lp[0]), float(li[0])

    def min_spacing(self):
        return lib().orc_min_spacing(self.h)

    def timestep(self, cfl):
        return lib().orc_timestep(self.h, float(cfl))

    def rk4(self, y, t0, dt, nsteps):
        y = np.array(y, dtype=np.float64, copy=True)
        lib().orc_rk4(self.h, _ptr(y), float(t0), float(dt), int(nsteps))
        return y


def annulus_errors(P, E, roots=(2, 1), r0=1.0, r1=2.0, m=1, k=None, cfl=0.5, tend=None,
                   truncate=True, nsteps=None):
    """The oracle's annulus_errors (annulus.jl:137-162): k the wavenumber of the
    mode (the caller's, from annulus_wavenumber)."""
    pr = Problem.annulus(P, E, roots=roots, r0=r0, r1=r1, truncate=truncate)
    t_end = tend if tend is not None else math.pi / 2 * max(m, 1) / k
    y0 = pr.truncate(pr.fill_exact(4, 0.0, (m, k)))
    if nsteps is None:
        nsteps = math.ceil(t_end / pr.timestep(cfl))
    dt = t_end / nsteps
    y = pr.rk4(y0, 0.0, dt, nsteps)
    err = y - pr.fill_exact(4, t_end, (m, k))
    l2, linf = pr.norms(err)
    l2u, _ = pr.norms(err, vars=(1,))
    l2v, _ = pr.norms(err, vars=(2,))
    h = pr.min_spacing() * (r1 - r0) / roots[1]
    return dict(l2=l2, l2u=l2u, l2v=l2v, linf=linf, h=h, nsteps=nsteps, dt=dt, tend=t_end,
                nblocks=pr.nb, energy0=pr.energy(y0), energy=pr.energy(y), y=y, y0=y0,
                problem=pr)


def nested_cubes(D, w0, levels):
    """Regions of nested cubes (sinewave.jl:192-199)."""
    c = 2 ** (levels - 1) * w0
    return [tuple((c - 2 ** (levels - 1 - k) * w0, c + 2 ** (levels - 1 - k) * w0)
                  for _ in range(D)) for k in range(1, levels)]


def wave_errors(D, P, E, roots=4, L=1.0, m=1, cfl=0.5, periods=0.25, truncate=True,
                reflecting=False, regions=(), nsteps=None):
    """The oracle's wave_errors (sinewave.jl:228-268)."""
    pr = Problem.brick(D, P, E, roots=roots, L=L, reflecting=reflecting, truncate=truncate,
                       regions=regions)
    kind = 1 if reflecting else 0
    y0 = pr.truncate(pr.fill_exact(kind, 0.0, (L, m)))
    omega = (math.pi if reflecting else 2 * math.pi) * m * math.sqrt(D) / L
    t_end = periods * 2 * math.pi / omega
    if nsteps is None:
        nsteps = math.ceil(t_end / pr.timestep(cfl))
    dt = t_end / nsteps
    y = pr.rk4(y0, 0.0, dt, nsteps)
    exact = pr.fill_exact(kind, t_end, (L, m))
    err = y - exact
    l2, linf = pr.norms(err)
    l2u, _ = pr.norms(err, vars=(1,))
    l2v, _ = pr.norms(err, vars=(2,))
    h = pr.min_spacing()
    return dict(l2=l2, l2u=l2u, l2v=l2v, linf=linf, h=h, he=(P + 1) * h, nsteps=nsteps,
                dt=dt, t_end=t_end, nblocks=pr.nb, energy0=pr.energy(y0),
                energy=pr.energy(y), y=y, y0=y0, problem=pr)


def sevenpatch_errors(P, E, roots=(2, 2), Lc=1.0, R1=3.0, k=(1.1, -0.7, 0.5), phi=0.3,
                      cfl=0.5, tend=1.0, truncate=True, grading=0.0, nsteps=None, Rm=None):
    """The oracle's sevenpatch_errors (sevenpatch.jl:84-117); with Rm the
    thirteen-patch ball."""
    if Rm is None:
        pr = Problem.ball(P, E, roots=roots, Lc=Lc, R1=R1, grading=grading, truncate=truncate,
                          k=k, phi=phi)
    else:
        pr = Problem.ball_ext(P, E, roots=roots, Lc=Lc, Rm=Rm, R1=R1, truncate=truncate, k=k,
                              phi=phi)
    y0 = pr.truncate(pr.fill_exact(2, 0.0))
    if nsteps is None:
        nsteps = math.ceil(tend / pr.timestep(cfl))
    dt = tend / nsteps
    y = pr.rk4(y0, 0.0, dt, nsteps)
    err = y - pr.fill_exact(2, tend)
    l2, linf = pr.norms(err)
    l2u, _ = pr.norms(err, vars=(1,))
    l2v, _ = pr.norms(err, vars=(2,))
    h = pr.min_spacing() * 2 * Lc / roots[0]
    return dict(l2=l2, l2u=l2u, l2v=l2v, linf=linf, h=h, nsteps=nsteps, dt=dt, tend=tend,
                nblocks=pr.nb, energy0=pr.energy(y0, 0.0), energy=pr.energy(y, tend), y=y,
                y0=y0, problem=pr)
