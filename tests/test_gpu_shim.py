"""The Julia extension's ccall sequences, executed through ctypes.

Julia is not in this image, so the package extension written out in the
integration notes cannot run here.  These tests make the same C calls in the
same order, with the same slots, through raw ctypes, and check that they give
what the Python host (the executed binding) gives: wave_rhs!, truncate_state!,
wave_energy, wave_norm / quadrature_norm, oldg_timestep, wave_integrator with
step! and solve!, scaled_modes, and the body of wave_errors
(sinewave.jl:228-268) composed from them.  The CPU test checks that every
symbol the extension calls is declared in include/twoldg.h and exported.
"""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import twoldg as T
from twoldg import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STATE, DU, AUX, ARG = 0, 1, 2, 3


def _julia_block():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"```julia\n(.*?)```", text, re.S)
    assert m, "the integration notes hold the Julia extension"
    return m.group(1)


def test_every_symbol_the_extension_calls_is_declared_and_exported():
    called = set(re.findall(r"ccall\(\(:(twoldg_\w+), lib\)", _julia_block()))
    assert {"twoldg_problem_create", "twoldg_rhs", "twoldg_truncate", "twoldg_rk4",
            "twoldg_energy", "twoldg_norms", "twoldg_timestep", "twoldg_modes"} <= called
    header = open(os.path.join(ROOT, "include", "twoldg.h")).read()
    lib = ctypes.CDLL(_lib.LIBPATH)
    for name in sorted(called):
        assert re.search(rf"\b{name}\(", header), name
        assert hasattr(lib, name), name


def _desc(forest, P, truncate=True, precision=_lib.F64):
    d = _lib.Desc()
    d.abi_version = _lib.ABI_VERSION
    d.D, d.P, d.E = forest.D, P, forest.E
    d.precision = precision
    d.truncate = int(truncate)
    d.nblocks = forest.nblocks
    sp = forest.spacings()
    faces = forest.faces()             # what twoldg_faces builds from the leaves
    d.spacing = sp.ctypes.data_as(ctypes.c_void_p)
    d.faces = ctypes.cast(faces, ctypes.c_void_p)
    return d, (sp, faces)


class Handle:
    """handle(p) of the extension: one library problem, destroyed with it."""

    def __init__(self, forest, P, truncate=True):
        self.L = _lib.lib()
        d, self._keep = _desc(forest, P, truncate)
        h = ctypes.c_void_p()
        _lib.check(self.L.twoldg_problem_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h

    def upload(self, slot, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        _lib.check(self.L.twoldg_state_upload(self.h, slot, u.ctypes.data_as(ctypes.c_void_p)))

    def download(self, slot, n):
        out = np.empty(n)
        _lib.check(self.L.twoldg_state_download(self.h, slot,
                                                out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def close(self):
        self.L.twoldg_problem_destroy(self.h)


def _d(x=0.0):
    return ctypes.c_double(x)


@pytest.mark.gpu
def test_gpu_shim_sequences_match_the_host_binding(rough):
    forest = T.wave_forest(2, 4, 2, roots=4, refined=True)
    P = 4
    p = T.WaveProblem(forest)
    h = Handle(forest, P)
    L = h.L
    n = p.state_len
    u = rough(n, 21)

    # wave_rhs!(du, u, p, t): ARG in, DU out
    h.upload(ARG, u)
    _lib.check(L.twoldg_rhs(h.h, ARG, DU, 0.3))
    du = h.download(DU, n)
    want = np.empty(n)
    T.wave_rhs(want, u, p, 0.3)
    assert np.array_equal(du, want)

    # truncate_state!(u, p)
    h.upload(ARG, u)
    _lib.check(L.twoldg_truncate(h.h, ARG))
    ut = h.download(ARG, n)
    assert np.array_equal(ut, T.truncate_state(u.copy(), p))

    # wave_energy(p, u), wave_norm(p, u; p, vars), quadrature_norm
    h.upload(ARG, ut)
    E = _d()
    _lib.check(L.twoldg_energy(h.h, ARG, 0.0, ctypes.byref(E)))
    assert E.value == T.wave_energy(p, ut)
    for pn, vars_ in ((2, (1, 2)), (math.inf, (1, 2)), (2, (1,)), (3, (2,))):
        h.upload(ARG, ut)
        lp, li = _d(), _d()
        mask = sum(1 << (v - 1) for v in vars_)
        _lib.check(L.twoldg_norms(h.h, ARG, -1, mask, 2 if pn == math.inf else pn,
                                  ctypes.byref(lp), ctypes.byref(li)))
        got = li.value if pn == math.inf else lp.value
        assert got == T.wave_norm(p, ut, pn, vars_) == T.quadrature_norm(p, ut, pn, vars_)

    # oldg_timestep(p; cfl)
    dt = _d()
    _lib.check(L.twoldg_timestep(h.h, 0.5, ctypes.byref(dt)))
    assert dt.value == T.oldg_timestep(p, 0.5)

    # wave_integrator + step! (one at a time) and solve! (the rest at once)
    dtv = dt.value
    h.upload(STATE, ut)
    for i in range(3):
        _lib.check(L.twoldg_rk4(h.h, i * dtv, dtv, 1))
    _lib.check(L.twoldg_rk4(h.h, 3 * dtv, dtv, 5))
    got = h.download(STATE, n)
    integ = T.wave_integrator(p, ut, 0.0, 8 * dtv, dtv)
    integ.step(3)
    integ.solve()
    assert np.array_equal(got, integ.u)

    # scaled_modes: the state set's data into STATE, then twoldg_modes
    hm = Handle(forest, P, truncate=False)
    hm.upload(STATE, u)
    M = 3
    modes = np.arange(1, M + 1, dtype=np.int32)
    w = np.array([T.doublefactorial(m) for m in modes], dtype=np.float64)
    pw = np.array([M / m for m in modes])
    tau = np.empty(forest.nblocks * forest.E ** 2)
    _lib.check(L.twoldg_modes(hm.h, STATE, 0, M, modes.ctypes.data_as(ctypes.c_void_p),
                              w.ctypes.data_as(ctypes.c_void_p),
                              pw.ctypes.data_as(ctypes.c_void_p), 1 / T.doublefactorial(M),
                              tau.ctypes.data_as(ctypes.c_void_p)))
    assert np.array_equal(tau, T.modal_indicator(p, mode=M, u=u).reshape(-1, order="F"))
    for x in (h, hm):
        x.close()
    p.close()


@pytest.mark.gpu
def test_gpu_wave_errors_body_over_the_shim():
    # sinewave.jl:228-268, line by line, each call through the C ABI.
    D, P, E, roots, m, cfl, periods = 2, 4, 2, 8, 1, 0.5, 0.25
    forest = T.wave_forest(D, P, E, roots=roots)
    el = T.Element(P)
    h = Handle(forest, P)
    L = h.L
    n = forest.nblocks * 2 * forest.N ** D
    from twoldg.problem import fill_at_nodes_forest
    # u0 = initial_state(problem, wave_exact(D, L, m, 0))
    u0 = fill_at_nodes_forest(forest, T.wave_exact(D, 1.0, m, 0.0), el)
    h.upload(ARG, u0)
    _lib.check(L.twoldg_truncate(h.h, ARG))
    u0 = h.download(ARG, n)
    t_end = periods * 2 * math.pi / T.wave_omega(D, 1.0, m)
    dt = _d()
    _lib.check(L.twoldg_timestep(h.h, cfl, ctypes.byref(dt)))
    nsteps = math.ceil(t_end / dt.value)
    dtv = t_end / nsteps
    h.upload(STATE, u0)
    _lib.check(L.twoldg_rk4(h.h, 0.0, dtv, nsteps))
    u = h.download(STATE, n)
    uexact = fill_at_nodes_forest(forest, T.wave_exact(D, 1.0, m, t_end), el)
    err = u - uexact
    out = {}
    for key, pn in (("l2", 2), ("linf", math.inf)):
        h.upload(ARG, err)
        lp, li = _d(), _d()
        _lib.check(L.twoldg_norms(h.h, ARG, -1, 3, 2, ctypes.byref(lp), ctypes.byref(li)))
        out[key] = li.value if pn == math.inf else lp.value
    for key, y in (("energy0", u0), ("energy", u)):
        h.upload(ARG, y)
        e = _d()
        _lib.check(L.twoldg_energy(h.h, ARG, 0.0, ctypes.byref(e)))
        out[key] = e.value
    h.close()
    want = T.wave_errors(D, P, E, roots=roots)
    for key in out:
        assert out[key] == want[key], key
    assert nsteps == want["nsteps"]
    # the published row (CODE.md:4142, the 80^2 periodic forest at P = 4)
    assert math.isclose(out["l2"], 4.54e-5, rel_tol=5e-3)
