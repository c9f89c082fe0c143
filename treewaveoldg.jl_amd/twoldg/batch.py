"""Batches of independent 2D problems advanced together on one GPU.

The reference runs one problem per call (wave_errors, sinewave.jl:228-268);
here many uniform 2D problems of different element orders, sizes and time
steps advance in one launch per RK4 stage (a device-resident work list), each
step a hipGraph replay.  See twoldg_batch_* in include/twoldg.h.
"""
import ctypes
import math

import numpy as np

from . import _lib
from .element import Element
from .forest import wave_forest
from .problem import PRECISIONS, _ptr


class Batch:
    """Problems given as (forest, f, dt): f(x, var) the initial data at the
    nodes (x of shape (D, n)), dt the problem's time step."""

    def __init__(self, problems, precision="f32c", truncate=True):
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}, got {precision!r}")
        self.problems = problems
        self._keep = []
        descs = (ctypes.c_void_p * len(problems))()
        dts = np.array([p[2] for p in problems], dtype=np.float64)
        self.sizes = []
        for i, (forest, f, dt) in enumerate(problems):
            if forest.D != 2:
                raise ValueError("a batch takes two-dimensional problems")
            sp = forest.spacings()
            faces = forest.faces()
            d = _lib.Desc()
            d.abi_version = _lib.ABI_VERSION
            d.D, d.P, d.E = forest.D, forest.P, forest.E
            d.precision = PRECISIONS[precision]
            d.truncate = int(truncate)
            d.nblocks = forest.nblocks
            d.spacing = _ptr(sp)
            d.faces = ctypes.cast(faces, ctypes.c_void_p)
            self._keep += [sp, faces, d]
            descs[i] = ctypes.cast(ctypes.pointer(d), ctypes.c_void_p)
            self.sizes.append(forest.nblocks * 2 * forest.N ** 2)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().twoldg_batch_create(descs, len(problems), _ptr(dts),
                                                  ctypes.byref(h)))
        self._h = h
        self.dt = dts

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().twoldg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, i, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.size != self.sizes[i]:
            raise ValueError(f"problem {i}'s state has {self.sizes[i]} entries, got {u.size}")
        _lib.check(_lib.lib().twoldg_batch_upload(self._h, i, _ptr(u)))

    def download(self, i):
        out = np.empty(self.sizes[i])
        _lib.check(_lib.lib().twoldg_batch_download(self._h, i, _ptr(out)))
        return out

    def rk4(self, nsteps):
        """nsteps RK4 steps of every problem; returns the HIP-event time in ms."""
        ms = ctypes.c_double()
        _lib.check(_lib.lib().twoldg_batch_rk4(self._h, int(nsteps), ctypes.byref(ms)))
        return ms.value

    def info(self):
        nodes = ctypes.c_int64()
        wg = ctypes.c_int64()
        tiles = (ctypes.c_int32 * 8)()
        tms = (ctypes.c_double * 16)()
        _lib.check(_lib.lib().twoldg_batch_info(self._h, ctypes.byref(nodes), ctypes.byref(wg),
                                                tiles, tms))
        return dict(nodes=nodes.value, workgroups=wg.value,
                    tiles={P: tiles[P - 1] for P in range(1, 9) if tiles[P - 1]},
                    tune_ms={P: tms[P - 1] for P in range(1, 9) if tiles[P - 1]},
                    tune_full_ms={P: tms[8 + P - 1] for P in range(1, 9) if tiles[P - 1]})


def gaussian_pulses(n=512, seed=20261203, orders=(2, 3, 4, 5, 6), roots=(4, 8, 16), E=2,
                    cfl=0.5):
    """The batched workload of independent 2D periodic unit boxes: per problem P
    uniform in `orders`, roots uniform in `roots`, E elements per block, and the
    radial Gaussian pulse u = exp(-|x - c|^2 / w^2), v = 0, with c uniform in
    [0.25, 0.75]^2 and w uniform in [0.03, 0.1]; dt = cfl h / (P+1).  Drawn with
    numpy's default_rng(seed) in the order P, roots, c, w per problem."""
    from .drivers import gaussian_pulse
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        P = int(rng.choice(orders))
        r = int(rng.choice(roots))
        c = rng.uniform(0.25, 0.75, size=2)
        w = float(rng.uniform(0.03, 0.1))
        forest = wave_forest(2, P, E, roots=r)
        dt = cfl * forest.minimum_spacing() / (P + 1)
        out.append((forest, gaussian_pulse(c, w), dt, dict(P=P, roots=r, c=c, w=w)))
    return out


def initial_states(problems):
    """The initial states of (forest, f, dt, ...) problems on the host: f at the
    nodes (v = 0 for the pulses, which truncation leaves unchanged)."""
    from .problem import fill_at_nodes_forest
    return [fill_at_nodes_forest(p[0], p[1], Element(p[0].P)) for p in problems]
