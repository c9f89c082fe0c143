"""WaveProblem and the hot-path entry points, over the HIP library.

Mirrors TreeWaveOLDG.jl's Julia API (src/TreeWaveOLDG.jl:59-78):

  WaveProblem(forest, el; truncate, geometry, boundary)   evolution.jl:1043-1108
  wave_rhs(du, u, p, t)          wave_rhs!        evolution.jl:1367-1371
  truncate_state(u, p)           truncate_state!  evolution.jl:1383-1400
  truncate_stage(u, integ, p, t) truncate_stage!  evolution.jl:1427-1430
  initial_state(p, f)            initial_state    evolution.jl:1440-1445
  wave_energy(p, u, t)           wave_energy      evolution.jl:1472-1489
  wave_norm(p, u, p, vars)       wave_norm        evolution.jl:1498-1512
  quadrature_norm(p, u, p, vars) quadrature_norm  layout.jl:223-235
  oldg_timestep(p, cfl)          oldg_timestep    evolution.jl:1528-1531
  wave_integrator(p, u0, t0, t1, dt)  wave_integrator  stepping.jl:72-78
  node_coordinates / fill_at_nodes    layout.jl:59-93

Julia's `!` functions mutate their first argument; here the numpy arrays passed
in are written in place the same way.  State vectors are host numpy arrays in
the statearray layout (block, variable, node; stepping.jl:36-41); the
integrator keeps its state on the device between steps, in the problem's state
slot, which the pure entry points (wave_rhs, truncate_state, wave_energy, the
norms) never touch: they pass their argument through the problem's argument
slot.  Every computation runs in libtwoldg on the GPU: there is no host
fallback.
"""
import ctypes
import math
import weakref

import numpy as np

from . import _lib
from .element import Element


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


PRECISIONS = {"f64": _lib.F64, "f32c": _lib.F32C, "split": _lib.SPLIT}


class WaveProblem:
    """Everything the right-hand side and the truncation need, built once and
    held on the device.

    `precision`: "f64" (the reference's structure, FP64 throughout), "f32c"
    (FP32 stencil and storage, compensated FP32 hi+lo RK update, FP64
    reductions) or "split" (u and its stencil in FP64, v in FP32 with a
    compensated step update; the fused 3D stages only).  `geometry`: None (flat,
    logical coordinates), a diagonal metric (PolarAnnulus) or a full one
    (SevenPatchBall, ThirteenPatchBall; a full geometry whose `outer_faces` is
    False has no outer sphere and takes no boundary).  `boundary`: the data on a
    full geometry's outer sphere, a PlaneWave (Dirichlet) or a RadiativeBoundary.
    """

    def __init__(self, forest, el=None, truncate=True, precision="f64", geometry=None,
                 boundary=None, generic=False):
        if el is None:
            el = Element(forest.P)
        if el.P != forest.P:
            raise ValueError(
                f"this field set holds elements of order P = {forest.P}, and the element is of "
                f"order P = {el.P}: the nodes, the stencil and the quadrature weights all "
                f"follow from P, so the two must agree")
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}, got {precision!r}")
        geometry = geometry if geometry is not None else forest.geometry
        full = geometry is not None and getattr(geometry, "kind", "full") == "full"
        radiative = boundary is not None and hasattr(boundary, "kappa")
        if radiative and not full:
            raise ValueError(
                "a radiative boundary forms the normal flux from the Cartesian gradient, which "
                "a full geometry carries: pass a `FullGeometry` (`geometry`)")
        if boundary is not None and not full:
            raise ValueError(
                "Dirichlet data (`boundary`) are given in Cartesian components and taken by a "
                "full geometry only: without one the gradient sets hold logical components, "
                "and the walls are TreeAMR's reflecting faces")
        # a full geometry without outer faces (geometry_boundary false,
        # geometry.jl:61-84) takes no boundary data
        outer = full and getattr(geometry, "outer_faces", True)
        if outer and boundary is None:
            raise ValueError(
                "the geometry has outer faces, and no data were given for them: pass "
                "`boundary = PlaneWave(k, phi)`, the Dirichlet data on the outer sphere, or "
                "`boundary = RadiativeBoundary()`")
        self.forest = forest
        self.element = el
        self.truncate = bool(truncate)
        self.precision = precision
        self.geometry = geometry
        self.boundary = boundary
        self.D, self.P, self.E, self.N = forest.D, forest.P, forest.E, forest.N
        self.nn = self.N ** self.D
        self.nblocks = forest.nblocks
        self.spacings = forest.spacings()
        self._faces = forest.faces()
        self._keep = []
        desc = _lib.Desc()
        desc.abi_version = _lib.ABI_VERSION
        desc.D, desc.P, desc.E = self.D, self.P, self.E
        desc.precision = PRECISIONS[precision]
        desc.truncate = int(self.truncate)
        desc.nblocks = self.nblocks
        desc.spacing = _ptr(self.spacings)
        desc.faces = ctypes.cast(self._faces, ctypes.c_void_p)
        desc.flags = _lib.FLAG_GENERIC if generic else 0
        self.hspeed = 0.0
        if geometry is not None and not full:
            # a diagonal metric (the annulus): sqrt(g), A_d and per-line truncation pairs
            m = geometry.metric(forest, el)
            self.dcoef = m["dcoef"]
            self.hspeed = m["hspeed"]
            desc.dcoef = _ptr(self.dcoef)
            desc.hspeed = self.hspeed
            if self.truncate:
                self._tables = geometry.truncation_tables(forest, el)
                desc.lmode = _ptr(self._tables["lmode"])
                desc.ldual = _ptr(self._tables["ldual"])
        elif geometry is not None:
            m = geometry.metric(forest, el)
            self.coef, self.flat = m["coef"], m["flat"]
            self.hspeed = m["hspeed"]
            desc.coef = _ptr(self.coef)
            desc.flat = _ptr(self.flat)
            desc.hspeed = self.hspeed
            if self.truncate:
                t = geometry.truncation_tables(forest, el)
                self._tables = t
                if t["ntables"]:
                    desc.tslot = _ptr(t["slot"])
                    desc.ntables = t["ntables"]
                    desc.tmodes = _ptr(t["modes"])
                    desc.tweights = _ptr(t["weights"])
                    desc.tchol = _ptr(t["chol"])
                if "fslot" in t:       # the factored truncation of exactly radial shells
                    desc.fslot = _ptr(t["fslot"])
                    desc.nfactored = t["nfactored"]
                    desc.fradial = _ptr(t["fradial"])
                    desc.frmodes = _ptr(t["frmodes"])
                    desc.frduals = _ptr(t["frduals"])
                    desc.faslot = _ptr(t["faslot"])
                    desc.nangular = t["nangular"]
                    desc.famodes = _ptr(t["famodes"])
                    desc.faweights = _ptr(t["faweights"])
                    desc.fachol = _ptr(t["fachol"])
            if radiative:
                # the outer faces take the radiative ghosts (radiative_faces,
                # evolution.jl:1191-1203)
                for i in range(len(self._faces)):
                    if self._faces[i].kind == _lib.FACE_DIRICHLET:
                        self._faces[i].kind = _lib.FACE_RADIATIVE
                desc.rad_kappa = boundary.kappa
                desc.rad_invR = boundary.invR(geometry)
                self.rad_invR = desc.rad_invR
            elif boundary is not None:
                self.dirichlet_x = geometry.face_positions(forest, el)
                desc.dirichlet_x = _ptr(self.dirichlet_x)
                for c in range(3):
                    desc.wave_k[c] = boundary.k[c]
                desc.wave_phi = boundary.phi
        self._desc = desc
        handle = ctypes.c_void_p()
        _lib.check(_lib.lib().twoldg_problem_create(ctypes.byref(desc), ctypes.byref(handle)))
        self._h = handle
        self.state_len = self.nblocks * 2 * self.nn
        self._state_owner = None  # weak reference to the integrator holding the state slot

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().twoldg_problem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- transfers ---------------------------------------------------------------
    def _check(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.shape != (self.state_len,):
            raise ValueError(
                f"the state vector has {u.size} entries but the field set's blocks hold "
                f"{self.state_len}")
        return u

    def upload(self, u, slot=_lib.SLOT_STATE):
        u = self._check(u)
        _lib.check(_lib.lib().twoldg_state_upload(self._h, slot, _ptr(u)))

    def download(self, slot=_lib.SLOT_STATE, out=None):
        if out is None:
            out = np.empty(self.state_len)
        _lib.check(_lib.lib().twoldg_state_download(self._h, slot, _ptr(out)))
        return out

    def truncate_state(self):
        """Truncate the state the problem holds (its state slot) in place; nothing when
        the problem does not truncate (truncate_state!, evolution.jl:1383-1395)."""
        if self.truncate:
            _lib.check(_lib.lib().twoldg_truncate(self._h, _lib.SLOT_STATE))

    def poison(self):
        """Fill every working array of the problem (not its state slots) with NaN;
        every call after this returns what it would have without it (the
        reference's tests poison `p.ufs.work` and the gradient sets the same way,
        test/refined_tests.jl:145-161)."""
        _lib.check(_lib.lib().twoldg_problem_poison(self._h))

    def info(self):
        n = ctypes.c_int64()
        b = ctypes.c_int64()
        f = ctypes.c_int32()
        _lib.check(_lib.lib().twoldg_problem_info(self._h, ctypes.byref(n), ctypes.byref(b),
                                                  ctypes.byref(f)))
        return {"state_len": n.value, "device_bytes": b.value, "fused": bool(f.value),
                "stages": {0: "generic", 1: "fused3d", 2: "batch2d"}[f.value]}

    # -- positions -------------------------------------------------------------
    def node_coordinates(self):
        """The positions initial data are evaluated at, (nblocks, D, N^D): physical
        on a full geometry (the seven-patch ball's data take x); logical on a flat
        forest and on a diagonal geometry, whose data take each patch's logical
        position (the reference's WithPatch callbacks, annulus.jl:95-135)."""
        x = self.forest.node_logical(self.element.offsets)
        if self.geometry is not None and getattr(self.geometry, "kind", "full") == "full":
            x = self.geometry.physical(self.forest, x)
        return x


def node_coordinates(forest):
    """Logical node positions of a flat forest from the library's host-side
    twoldg_node_coordinates, (nblocks, D, N^D) (layout.jl:59-76)."""
    sp = forest.spacings()
    desc = _lib.Desc()
    desc.abi_version = _lib.ABI_VERSION
    desc.D, desc.P, desc.E = forest.D, forest.P, forest.E
    desc.precision = _lib.F64
    desc.nblocks = forest.nblocks
    desc.spacing = _ptr(sp)
    lo = np.ascontiguousarray([forest.patches[p][1][:forest.D] for (p, lev, *c) in forest.leaves],
                              dtype=np.float64).reshape(forest.nblocks, forest.D)
    corner = np.ascontiguousarray([c[:forest.D] for (p, lev, *c) in forest.leaves],
                                  dtype=np.int64).reshape(forest.nblocks, forest.D)
    x = np.empty((forest.nblocks, forest.D, forest.N ** forest.D))
    _lib.check(_lib.lib().twoldg_node_coordinates(ctypes.byref(desc), _ptr(lo),
                                                  corner.ctypes.data_as(ctypes.c_void_p),
                                                  _ptr(x)))
    return x


def exchange_plan(forest, el=None):
    """The exchange operator the library builds for `forest` (host only): a dict
    of rowptr, cb, cl, vu, vg (see include/twoldg.h twoldg_exchange_plan)."""
    faces = forest.faces()
    sp = forest.spacings()
    desc = _lib.Desc()
    desc.abi_version = _lib.ABI_VERSION
    desc.D, desc.P, desc.E = forest.D, forest.P, forest.E
    desc.precision = _lib.F64
    desc.truncate = 1
    desc.nblocks = forest.nblocks
    desc.spacing = _ptr(sp)
    desc.faces = ctypes.cast(faces, ctypes.c_void_p)
    dummy = None
    if forest.geometry is not None:
        # the plan does not read them, but the description must be complete
        dummy = np.zeros(1)
        desc.coef = _ptr(dummy)
        desc.dirichlet_x = _ptr(dummy)
    L = _lib.lib()
    nrows = ctypes.c_int64()
    nnz = ctypes.c_int64()
    _lib.check(L.twoldg_exchange_plan(ctypes.byref(desc), ctypes.byref(nrows), ctypes.byref(nnz),
                                      None, None, None, None, None))
    rowptr = np.zeros(nrows.value + 1, dtype=np.int32)
    cb = np.zeros(nnz.value, dtype=np.int32)
    cl = np.zeros(nnz.value, dtype=np.int32)
    vu = np.zeros(nnz.value)
    vg = np.zeros(nnz.value)
    _lib.check(L.twoldg_exchange_plan(ctypes.byref(desc), None, None, _ptr(rowptr), _ptr(cb),
                                      _ptr(cl), _ptr(vu), _ptr(vg)))
    return dict(rowptr=rowptr, cb=cb, cl=cl, vu=vu, vg=vg)


# ---------------------------------------------------------------------------
# The reference's entry points.
# ---------------------------------------------------------------------------
def wave_rhs(du, u, p, t=0.0):
    """du <- f(u, t) in place (wave_rhs!).  Pure: u is not written, and an
    integrator of p keeps its state."""
    p.upload(u, _lib.SLOT_ARG)
    _lib.check(_lib.lib().twoldg_rhs(p._h, _lib.SLOT_ARG, _lib.SLOT_DU, float(t)))
    p.download(_lib.SLOT_DU, out=du)
    return None


def truncate_state(u, p):
    """v <- (Pi x ... x Pi) v in place; nothing when p does not truncate."""
    if not p.truncate:
        return u
    p.upload(u, _lib.SLOT_ARG)
    _lib.check(_lib.lib().twoldg_truncate(p._h, _lib.SLOT_ARG))
    p.download(_lib.SLOT_ARG, out=u)
    return u


def truncate_stage(u, integrator, p, t):
    """truncate_state as the integrator's limiter hook."""
    truncate_state(u, p)
    return None


def _fill(x, f):
    nb, D, nn = x.shape
    u = np.empty((nb, 2, nn))
    xs = np.moveaxis(x, 1, 0).reshape(D, nb * nn)
    for var in (1, 2):
        u[:, var - 1, :] = np.asarray(f(xs, var), dtype=np.float64).reshape(nb, nn)
    return u.reshape(-1)


def fill_at_nodes(f, p):
    """The state vector of f(x, var) at every owned node (var = 1 for u, 2 for
    v); f takes x as an array (D, n) of positions and returns n values."""
    return _fill(p.node_coordinates(), f)


def fill_at_nodes_forest(forest, f, el):
    """fill_at_nodes on a flat forest, without a problem."""
    return _fill(forest.node_logical(el.offsets), f)


def initial_state(p, f):
    """f at the nodes, with v truncated when p truncates."""
    u0 = fill_at_nodes(f, p)
    return truncate_state(u0, p)


def wave_energy(p, u, t=None):
    """The discrete energy 1/2 sum w (v^2 + sum_d (D_d u)^2), on a full geometry
    weighted by sqrt(g) with the Cartesian gradient."""
    if p.boundary is not None and not hasattr(p.boundary, "kappa") and t is None:
        raise ValueError("the problem has Dirichlet data, which depend on the time, and none "
                         "was given: pass `t`")
    p.upload(u, _lib.SLOT_ARG)
    out = ctypes.c_double()
    _lib.check(_lib.lib().twoldg_energy(p._h, _lib.SLOT_ARG, float(t or 0.0), ctypes.byref(out)))
    return out.value


def _norm(p, u, pn, vars):
    mask = 0
    for v in vars:
        if v not in (1, 2):
            raise ValueError(f"variables are 1 (u) and 2 (v), got {v}")
        mask |= 1 << (v - 1)
    if not (pn == math.inf or (float(pn).is_integer() and pn >= 1)):
        raise ValueError(f"quadrature_norm takes an integer p >= 1 or Inf, got p = {pn}")
    p.upload(u, _lib.SLOT_ARG)
    lp = ctypes.c_double()
    li = ctypes.c_double()
    _lib.check(_lib.lib().twoldg_norms(p._h, _lib.SLOT_ARG, -1, mask,
                                       2 if pn == math.inf else int(pn), ctypes.byref(lp),
                                       ctypes.byref(li)))
    return li.value if pn == math.inf else lp.value


def quadrature_norm(p, u, pn=2, vars=(1, 2)):
    """The p-norm under the Gauss-Lobatto quadrature normalized by the volume
    (and the number of variables); p = inf is the largest |value|."""
    return _norm(p, u, pn, vars)


def wave_norm(p, u, pn=2, vars=(1, 2)):
    """quadrature_norm, on a geometry weighted by sqrt(g)."""
    return _norm(p, u, pn, vars)


def oldg_timestep(p, cfl):
    """cfl * h / (P+1) at the smallest spacing, on a geometry the smallest
    h_b / speed_b."""
    out = ctypes.c_double()
    _lib.check(_lib.lib().twoldg_timestep(p._h, float(cfl), ctypes.byref(out)))
    return out.value


class Integrator:
    """Classic RK4 with truncation as stage and step limiter, state on the
    device (IMEXRungeKutta's init/step!/solve! semantics, stepping.jl:72-78).

    The integrator owns its problem's state slot: the problem's pure entry
    points leave it alone.  Building another integrator on the same problem
    takes the slot over, and this one then refuses to go on rather than step
    the other's state."""

    def __init__(self, p, u0, t0, t1, dt):
        if not dt > 0:
            raise ValueError("dt must be positive")
        self.p = p
        self.t0, self.t1 = float(t0), float(t1)
        # ceil((t1 - t0)/dt) equal steps (stepping.jl:55-61); a quotient within
        # a few ulp of a whole number m is m steps of the dt given
        span = self.t1 - self.t0
        r = span / dt
        m = round(r)
        if m >= 1 and abs(r - m) <= 4 * (math.ulp(r) + math.ulp(span) / dt):
            self.nsteps, self.dt = int(m), float(dt)
        else:
            self.nsteps = max(0, math.ceil(r))
            self.dt = span / self.nsteps if self.nsteps else float(dt)
        self.istep = 0
        p.upload(u0, _lib.SLOT_STATE)
        # weak: the problem must not keep its integrator alive (a cycle would
        # leave both, and the device memory, to the cyclic garbage collector)
        p._state_owner = weakref.ref(self)

    def _own(self):
        owner = self.p._state_owner
        if owner is None or owner() is not self:
            raise RuntimeError(
                "another integrator was built on this problem since, and holds its state "
                "slot: build one problem per concurrent integrator")

    @property
    def t(self):
        return self.t0 + self.istep * self.dt

    @property
    def u(self):
        self._own()
        return self.p.download(_lib.SLOT_STATE)

    def timed(self, n):
        """n steps timed with HIP events on the problem's stream: (total ms,
        summed stage-kernel ms, stage launches)."""
        self._own()
        ms = ctypes.c_double()
        kms = ctypes.c_double()
        nl = ctypes.c_int64()
        _lib.check(_lib.lib().twoldg_rk4_timed(self.p._h, self.t, self.dt, int(n), 0,
                                               ctypes.byref(ms), ctypes.byref(kms),
                                               ctypes.byref(nl)))
        self.istep += n
        return ms.value, kms.value, nl.value

    def step(self, n=1):
        self._own()
        n = min(n, self.nsteps - self.istep)
        if n <= 0:
            return
        _lib.check(_lib.lib().twoldg_rk4(self.p._h, self.t, self.dt, n))
        self.istep += n

    def solve(self):
        self.step(self.nsteps - self.istep)
        return self

    def energy(self):
        self._own()
        out = ctypes.c_double()
        _lib.check(_lib.lib().twoldg_energy(self.p._h, _lib.SLOT_STATE, self.t,
                                            ctypes.byref(out)))
        return out.value


def wave_integrator(p, u0, t0, t1, dt):
    return Integrator(p, u0, t0, t1, dt)
