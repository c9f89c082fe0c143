"""Problems and convergence drivers (TreeWaveOLDG.jl src/sinewave.jl, src/sevenpatch.jl).

wave_omega / wave_exact        sinewave.jl:17-39
reflect_omega / reflect_exact  sinewave.jl:47-76
wave_errors / reflect_errors   sinewave.jl:228-274
PlaneWave / sevenpatch_exact   sevenpatch.jl:19-59
sevenpatch_errors              sevenpatch.jl:84-117
convergence_rate               evolution.jl:1534-1540

The exact solutions are evaluated on the host (they are the caller's initial
and comparison data, Julia closures in the reference); everything between the
initial state and the error norms runs in libtwoldg.
"""
import math

import numpy as np

from .element import Element
from .forest import wave_forest
from .problem import (WaveProblem, initial_state, fill_at_nodes, oldg_timestep, wave_integrator,
                      wave_norm, wave_energy, _lib)


def convergence_rate(hs, errs):
    """The least-squares slope of log(errs) against log(hs): the convergence
    rate of a sequence of runs (evolution.jl:1534-1540)."""
    x = np.log(np.asarray(hs, dtype=np.float64))
    y = np.log(np.asarray(errs, dtype=np.float64))
    xm, ym = x.mean(), y.mean()
    return float(np.sum((x - xm) * (y - ym)) / np.sum((x - xm) ** 2))


def wave_omega(D, L, m):
    return 2 * math.pi * m * math.sqrt(D) / L


def wave_exact(D, L, m, t):
    """u = cos(w t) prod sin(2 pi m x_d / L), v = du/dt."""
    w = wave_omega(D, L, m)

    def f(x, var):
        shape = np.ones(x.shape[1])
        for d in range(D):
            shape = shape * np.sin(2 * math.pi * m * x[d] / L)
        return math.cos(w * t) * shape if var == 1 else -w * math.sin(w * t) * shape
    return f


def reflect_omega(D, L, m):
    return math.pi * m * math.sqrt(D) / L


def _cospi(z):
    r = np.fmod(np.abs(z), 2.0)
    r = np.where(r > 1, 2 - r, r)
    return np.where(r <= 0.25, np.cos(math.pi * r),
                    np.where(r <= 0.75, np.sin(math.pi * (0.5 - r)), -np.cos(math.pi * (1 - r))))


def reflect_exact(D, L, m, t):
    """u = cos(w t) prod cos(m pi x_d / L), the Neumann standing mode."""
    w = reflect_omega(D, L, m)

    def f(x, var):
        shape = np.ones(x.shape[1])
        for d in range(D):
            shape = shape * _cospi(m * x[d] / L)
        return math.cos(w * t) * shape if var == 1 else -w * math.sin(w * t) * shape
    return f


def gaussian_pulse(c, w):
    """u = exp(-|x - c|^2 / w^2), v = 0."""
    c = np.asarray(c, dtype=np.float64)

    def f(x, var):
        if var == 2:
            return np.zeros(x.shape[1])
        r2 = sum((x[d] - c[d]) ** 2 for d in range(len(c)))
        return np.exp(-r2 / (w * w))
    return f


class PlaneWave:
    """u = sin(k.x - |k| t + phi) (sevenpatch.jl:19-36)."""

    def __init__(self, k, phi):
        self.k = tuple(float(x) for x in k)
        self.omega = math.sqrt(sum(x * x for x in self.k))
        self.phi = float(phi)

    def phase(self, x, t):
        s = self.k[0] * x[0]
        for d in range(1, len(self.k)):
            s = s + self.k[d] * x[d]
        return s - self.omega * t + self.phi

    def __call__(self, x, t):
        th = self.phase(x, t)
        c = np.cos(th)
        return (np.sin(th),) + tuple(k * c for k in self.k)


def sevenpatch_exact(wave, t):
    def f(x, var):
        th = wave.phase(x, t)
        return np.sin(th) if var == 1 else -wave.omega * np.cos(th)
    return f


def wave_errors(D, P, E, roots=4, L=1.0, m=1, cfl=0.5, periods=0.25, truncate=True,
                reflecting=False, refined=False, precision="f64", nsteps=None):
    """Evolve the standing mode to a quarter period (by default) and return the
    L2 and L-inf errors of u and v together, the spacing, the step count and the
    energies (sinewave.jl:228-268)."""
    forest = wave_forest(D, P, E, roots=roots, L=L, reflecting=reflecting, refined=refined)
    el = Element(P)
    problem = WaveProblem(forest, el, truncate=truncate, precision=precision)
    exact_mode = reflect_exact if reflecting else wave_exact
    omega = reflect_omega if reflecting else wave_omega
    u0 = initial_state(problem, exact_mode(D, L, m, 0.0))
    t_end = periods * 2 * math.pi / omega(D, L, m)
    if nsteps is None:
        nsteps = math.ceil(t_end / oldg_timestep(problem, cfl))
    dt = t_end / nsteps
    integ = wave_integrator(problem, u0, 0.0, t_end, dt)
    integ.solve()
    energy = integ.energy()
    uexact = fill_at_nodes(exact_mode(D, L, m, t_end), problem)
    problem.upload(uexact, _lib.SLOT_AUX)
    l2, linf = _errors(problem, (1, 2))
    l2u, _ = _errors(problem, (1,))
    l2v, _ = _errors(problem, (2,))
    h = forest.minimum_spacing()
    return dict(l2=l2, l2u=l2u, l2v=l2v, linf=linf, h=h, he=(P + 1) * h, nsteps=nsteps, dt=dt,
                t_end=t_end, nblocks=forest.nblocks, energy0=wave_energy(problem, u0),
                energy=energy, problem=problem, integrator=integ)


def _errors(problem, vars):
    """Norms of state - exact (exact in the aux slot), on the device."""
    import ctypes
    mask = sum(1 << (v - 1) for v in vars)
    lp = ctypes.c_double()
    li = ctypes.c_double()
    _lib.check(_lib.lib().twoldg_norms(problem._h, _lib.SLOT_STATE, _lib.SLOT_AUX, mask, 2,
                                       ctypes.byref(lp), ctypes.byref(li)))
    return lp.value, li.value


def reflect_errors(D, P, E, **kw):
    """wave_errors on a reflecting box (sinewave.jl:271-274)."""
    kw.pop("reflecting", None)
    return wave_errors(D, P, E, reflecting=True, **kw)


def ball_forest(P, E, roots=(2, 2), Lc=1.0, Rm=None, R1=3.0, refined=False, grading=0.0):
    """The seven-patch ball, or with a middle radius Rm the thirteen-patch one
    (roots (n, m, m')), whose shells are not graded (sevenpatch.jl:119-129)."""
    if Rm is None:
        from .geometry import sevenpatch_forest
        return sevenpatch_forest(P, E, roots=roots, Lc=Lc, R1=R1, refined=refined,
                                 grading=grading)
    if grading != 0:
        raise ValueError(f"the thirteen-patch ball's shells are not graded, got grading = {grading}")
    from .thirteenpatch import thirteenpatch_forest
    return thirteenpatch_forest(P, E, roots=roots, Lc=Lc, Rm=Rm, R1=R1, refined=refined)


def sevenpatch_errors(P, E, roots=(2, 2), Lc=1.0, R1=3.0, k=(1.1, -0.7, 0.5), phi=0.3,
                      cfl=0.5, tend=1.0, truncate=True, grading=0.0, precision="f64",
                      nsteps=None, Rm=None, refined=False):
    """A plane wave through the seven-patch ball (or, with Rm, the thirteen-patch
    ball) with the wave itself as the Dirichlet data on the outer sphere
    (sevenpatch.jl:84-117)."""
    forest, geo = ball_forest(P, E, roots=roots, Lc=Lc, Rm=Rm, R1=R1, refined=refined,
                              grading=grading)
    el = Element(P)
    wave = PlaneWave(k, phi)
    problem = WaveProblem(forest, el, truncate=truncate, precision=precision, geometry=geo,
                          boundary=wave)
    u0 = initial_state(problem, sevenpatch_exact(wave, 0.0))
    if nsteps is None:
        nsteps = math.ceil(tend / oldg_timestep(problem, cfl))
    dt = tend / nsteps
    energy0 = wave_energy(problem, u0, t=0.0)
    integ = wave_integrator(problem, u0, 0.0, tend, dt)
    integ.solve()
    energy = integ.energy()
    uexact = fill_at_nodes(sevenpatch_exact(wave, tend), problem)
    problem.upload(uexact, _lib.SLOT_AUX)
    l2, linf = _errors(problem, (1, 2))
    l2u, _ = _errors(problem, (1,))
    l2v, _ = _errors(problem, (2,))
    h = forest.minimum_spacing() * 2 * Lc / roots[0]
    return dict(l2=l2, l2u=l2u, l2v=l2v, linf=linf, h=h, he=(P + 1) * h, nsteps=nsteps, dt=dt,
                tend=tend, nblocks=forest.nblocks, energy0=energy0, energy=energy,
                problem=problem, integrator=integ)


def radiative_errors(P, E, roots=(2, 2), Lc=1.0, R1=3.0, w=0.5, kappa=1.0, inverse_r=True,
                     cfl=0.5, tend=2.0, truncate=True, refined=False, grading=0.0, Rm=None,
                     precision="f64", nsteps=None):
    """An OutgoingPulse of width w through a RadiativeBoundary on the seven-patch
    ball (or, with Rm, the thirteen-patch one) to tend: the sqrt(g)-weighted L2
    errors together and apart, the L-inf error, the cube's spacing, the steps and
    the energies, the boundary's own included (sevenpatch.jl:187-236)."""
    from .thirteenpatch import OutgoingPulse, RadiativeBoundary, pulse_state
    forest, geo = ball_forest(P, E, roots=roots, Lc=Lc, Rm=Rm, R1=R1, refined=refined,
                              grading=grading)
    el = Element(P)
    pulse = OutgoingPulse(w)
    problem = WaveProblem(forest, el, truncate=truncate, precision=precision, geometry=geo,
                          boundary=RadiativeBoundary(kappa, inverse_r))
    u0 = initial_state(problem, pulse_state(pulse, 0.0))
    if nsteps is None:
        nsteps = math.ceil(tend / oldg_timestep(problem, cfl))
    dt = tend / nsteps
    energy0 = wave_energy(problem, u0)
    integ = wave_integrator(problem, u0, 0.0, tend, dt)
    integ.solve()
    energy = integ.energy()
    uexact = fill_at_nodes(pulse_state(pulse, tend), problem)
    problem.upload(uexact, _lib.SLOT_AUX)
    l2, linf = _errors(problem, (1, 2))
    l2u, _ = _errors(problem, (1,))
    l2v, _ = _errors(problem, (2,))
    h = forest.minimum_spacing() * 2 * Lc / roots[0]
    return dict(l2=l2, l2u=l2u, l2v=l2v, linf=linf, h=h, he=(P + 1) * h, nsteps=nsteps, dt=dt,
                tend=tend, nblocks=forest.nblocks, energy0=energy0, energy=energy,
                problem=problem, integrator=integ)
