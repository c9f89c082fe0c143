"""The 2D blast wave and the adaptive run that follows it.

Mirrors TreeWaveOLDG.jl src/blast.jl and the chunk helpers of
src/precision.jl:33-70:

  blast_initial(L, x0, sigma)          blast.jl:14-35
  blast_reference(L, x0, sigma; rmax)  blast.jl:37-70
  blast_radial_table(ref, t)           blast.jl:72-98
  blast_exact(ref, t)                  blast.jl:100-131
  uniform_blast(...)                   blast.jl:133-195
  adaptive_blast(...)                  blast.jl:197-360 (the regrid cycle 336-348)
  chunk_count, chunk_bounds, wrap      precision.jl:33-70

The exact solution is host work (a Hankel-transform table built once, as in
the reference); the evolution, the indicator and the regrid transfer run on
the GPU.
"""
import math

import numpy as np
from scipy.special import j0

from .element import Element
from .forest import wave_forest
from .problem import (WaveProblem, initial_state, fill_at_nodes, oldg_timestep, wave_integrator,
                      wave_energy, quadrature_norm)
from .regrid import (check_thresholds, check_mode, refine_flags, refinement_buffer,
                     adapt_to_initial_data, regrid)


def wrap(x, L):
    """x reduced into [0, L), as x - L floor(x/L) (precision.jl:62-68)."""
    return x - L * np.floor(x / L)


def chunk_count(t_end, chunk):
    """ceil(t_end/chunk), except that a quotient within a few ulp of a whole m
    gives m (precision.jl:33-46)."""
    r = t_end / chunk
    m = round(r)
    tol = 4 * (math.ulp(r) + math.ulp(t_end) / chunk)
    return int(m) if m >= 1 and abs(r - m) <= tol else math.ceil(r)


def chunk_bounds(c, nchunks, t_end, chunk):
    """The span of chunk c (1-based) of chunk_count(t_end, chunk) chunks, the last
    ending at t_end exactly (precision.jl:48-60)."""
    tstart = min((c - 1) * chunk, t_end)
    t = t_end if c == nchunks else min(c * chunk, t_end)
    return tstart, t


def blast_initial(L, x0, sigma):
    """A radial Gaussian at rest, u = exp(-(r/sigma)^2), r the periodically wrapped
    distance from x0 in a box of side L (blast.jl:14-35)."""

    def f(x, var):
        if var != 1:
            return np.zeros(x.shape[1])
        r2 = np.zeros(x.shape[1])
        for d in range(2):
            r2 += (wrap(x[d] - x0[d] + L / 2, L) - L / 2) ** 2
        return np.exp(-((np.sqrt(r2) / sigma) ** 2))

    return f


def blast_reference(L, x0, sigma, *, rmax, nr=2001, nk=2000):
    """The time-independent part of the exact solution: the inverse Hankel
    transform of the Gaussian, f^(k) = (sigma^2/2) exp(-k^2 sigma^2/4), cut at
    k = 13/sigma and taken by the midpoint rule, with J0(k r) tabulated on the
    (r, k) grid (blast.jl:37-70)."""
    dk = (13 / sigma) / nk
    ks = (np.arange(1, nk + 1) - 0.5) * dk
    ws = (sigma ** 2 / 2) * np.exp(-(ks * sigma) ** 2 / 4) * ks * dk
    rs = np.linspace(0.0, rmax, nr)
    J = j0(np.multiply.outer(rs, ks))
    return dict(L=float(L), x0=tuple(float(a) for a in x0), sigma=float(sigma), rs=rs,
                dr=rs[1] - rs[0], rmax=float(rmax), ks=ks, ws=ws, J=J)


def blast_radial_table(ref, t):
    """The free-space radial profiles of u and du/dt at time t on the radius grid:
    u = sum_k w_k J0(k r) cos(k t), du/dt = -sum_k w_k k J0(k r) sin(k t)."""
    ks, ws = ref["ks"], ref["ws"]
    us = ref["J"] @ (ws * np.cos(ks * t))
    vs = -(ref["J"] @ (ws * np.sin(ks * t) * ks))
    return us, vs


def blast_exact(ref, t):
    """The exact solution at time t as an (x, var) callback: the radial profile
    summed over the periodic images of the source out to rmax, linear in r
    (blast.jl:100-131)."""
    us, vs = blast_radial_table(ref, float(t))
    L, rmax, dr = ref["L"], ref["rmax"], ref["dr"]
    x0 = ref["x0"]
    nim = math.ceil(rmax / L)

    def f(x, var):
        tab = us if var == 1 else vs
        out = np.zeros(x.shape[1])
        for j1 in range(-nim, nim + 1):
            for j2 in range(-nim, nim + 1):
                r = np.hypot(x[0] - x0[0] - j1 * L, x[1] - x0[1] - j2 * L)
                ok = r < rmax
                q = r[ok] / dr
                i = np.floor(q).astype(np.int64)
                s = q - i
                out[ok] += (1 - s) * tab[i] + s * tab[i + 1]
        return out

    return f


def _errors(p, u, ue):
    err = u - ue
    return quadrature_norm(p, err, pn=math.inf), quadrature_norm(p, err)


def uniform_blast(*, P, E, roots, L=1.0, sigma=2 / 25, t_end=2 / 5, cfl=0.5, truncate=True,
                  precision="f64"):
    """The blast wave on a uniform roots x roots box; L-inf and L2 errors at t_end
    against blast_exact, the spacing, the node count, the steps and the energy at
    the start and the end (blast.jl:133-195)."""
    forest = wave_forest(2, P, E, roots=roots, L=L)
    el = Element(P)
    p = WaveProblem(forest, el, truncate=truncate, precision=precision)
    x0 = (L / 2, L / 2)
    u0 = initial_state(p, blast_initial(L, x0, sigma))
    ref = blast_reference(L, x0, sigma, rmax=t_end + 6 * sigma)
    nsteps = math.ceil(t_end / oldg_timestep(p, cfl))
    dt = t_end / nsteps
    integ = wave_integrator(p, u0, 0.0, t_end, dt)
    integ.solve()
    u = integ.u
    linf, l2 = _errors(p, u, fill_at_nodes(blast_exact(ref, t_end), p))
    out = dict(err=linf, l2=l2, h=forest.minimum_spacing(), nodes=forest.nblocks * forest.N ** 2,
               nsteps=nsteps, energy0=wave_energy(p, u0), energy=wave_energy(p, u))
    p.close()
    return out


def adaptive_blast(*, P, E, roots, maxlevel_cap, refine_tol, coarsen_tol, chunk, buffer=None,
                   L=1.0, sigma=2 / 25, t_end=2 / 5, cfl=0.5, truncate=True, maxpasses=8,
                   precision="f64"):
    """The blast wave on a forest that follows it (blast.jl:197-360): the initial
    hierarchy from adapt_to_initial_data, then chunks of `chunk` in time, each a
    fresh integrator at the step of the current finest spacing, and after every
    chunk but the last a regrid by refine_flags of u's mode P - truncate, with the
    state truncated again on the new mesh (coarsening brings v's top mode back).

    Returns the L-inf and L2 errors at t_end against blast_exact, the finest
    spacing, the final node count, the largest at any time, the node-steps, the
    steps, the regrids that changed the mesh, the final maxlevel, the forest, the
    adaptation's passes and whether it converged, the energy at the start and
    the end, and one record per chunk."""
    check_thresholds(refine_tol, coarsen_tol)
    if not chunk > 0:
        raise ValueError(f"chunk = {chunk}: the time between regrids must be positive")
    mode = P - int(truncate)
    check_mode(mode, P)
    forest = wave_forest(2, P, E, roots=roots, L=L)
    el = Element(P)
    x0 = (L / 2, L / 2)
    initial = blast_initial(L, x0, sigma)
    ref = blast_reference(L, x0, sigma, rmax=t_end + 6 * sigma)
    if buffer is None:
        buffer = refinement_buffer(forest, maxlevel_cap, chunk)

    def flags(q):
        return refine_flags(q, el, mode=mode, refine_tol=refine_tol, coarsen_tol=coarsen_tol,
                            maxlevel_cap=maxlevel_cap)

    forest, passes, converged = adapt_to_initial_data(
        forest, el, initial=initial, flags=flags, buffer=buffer, maxpasses=maxpasses,
        precision=precision)
    p = WaveProblem(forest, el, truncate=truncate, precision=precision)
    u = initial_state(p, initial)
    energy0 = wave_energy(p, u)

    def nodes(f):
        return f.nblocks * f.N ** 2

    nodes_max = nodes(forest)
    node_steps = nsteps_total = nregrids = 0
    chunks = []
    nchunks = chunk_count(t_end, chunk)
    t = 0.0
    for c in range(1, nchunks + 1):
        tstart, t = chunk_bounds(c, nchunks, t_end, chunk)
        if not t > tstart:
            break
        e_start = wave_energy(p, u)
        dt = oldg_timestep(p, cfl)
        integ = wave_integrator(p, u, tstart, t, dt)
        integ.solve()
        u = integ.u
        e_end = wave_energy(p, u)
        nsteps_total += integ.nsteps
        node_steps += integ.nsteps * nodes(p.forest)
        record = dict(tstart=tstart, t=t, nsteps=integ.nsteps, nblocks=p.forest.nblocks,
                      maxlevel=p.forest.maxlevel(), energy_start=e_start, energy_end=e_end)
        # no regrid after the last chunk: the errors are measured on that mesh
        changed = False
        if c < nchunks:
            q, changed = regrid(p, flags(p), buffer)
            if changed:
                nregrids += 1
                q.truncate_state()
                u = q.download()
                p.close()
                p = q
                nodes_max = max(nodes_max, nodes(p.forest))
        record["regridded"] = changed
        chunks.append(record)

    linf, l2 = _errors(p, u, fill_at_nodes(blast_exact(ref, t_end), p))
    out = dict(err=linf, l2=l2, h=p.forest.minimum_spacing(), nodes=nodes(p.forest),
               nodes_max=nodes_max, node_steps=node_steps, nsteps=nsteps_total,
               nregrids=nregrids, maxlevel=p.forest.maxlevel(), forest=p.forest, passes=passes,
               converged=converged, energy0=energy0, energy=wave_energy(p, u), chunks=chunks)
    p.close()
    return out
