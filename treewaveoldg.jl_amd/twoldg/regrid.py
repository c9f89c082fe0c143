"""Adaptive refinement: the modal indicator, the block verdict and the regrid of a
problem's state.

Mirrors TreeWaveOLDG.jl's src/refinement.jl and the regrid cycle of
src/blast.jl:336-348:

  modal_indicator(p, el; mode, var)    refinement.jl:148-155 (mode_kernel! 56-80)
  mode_amplitude(p, el; mode, var)     refinement.jl:168-174
  refine_mark(tau, el, b, k; ...)      refinement.jl:227-253
  refine_flags(p, el; ...)             refinement.jl:279-287
  refinement_buffer(forest, cap, d)    refinement.jl:306-316
  doublefactorial, check_mode, check_thresholds   refinement.jl:178-188, 255-261

and, over the HIP library's twoldg_transfer, the part of TreeAMR's regrid!
that moves a state (refinement by the coarse element's polynomial, coarsening
by the exact L2 projection, both element-local), with the forest logic of
regrid! restated from what the reference documents of it: per-block flags
Refine / Keep / Coarsen, a flag's box dilated by the buffer recruits the blocks
it reaches to the flagging block's target level, 2^D siblings coarsen only
together and only where the result stays 2:1 balanced across faces, and the
forest is balanced after.  TreeAMR itself is not vendored with the reference,
so these mesh decisions follow its documented behaviour rather than its code.

The indicator and the amplitudes run on the GPU (twoldg_modes); the verdict and
the forest logic are host work on per-block values, as in the reference, where
they are TreeAMR's callbacks.  The state moves on the GPU (twoldg_transfer).
"""
import ctypes
import enum
import itertools
import math

import numpy as np

from . import _lib
from .problem import WaveProblem, fill_at_nodes


class RegridFlag(enum.IntEnum):
    """TreeAMR's per-block regrid flags."""
    Coarsen = -1
    Keep = 0
    Refine = 1


Coarsen, Keep, Refine = RegridFlag.Coarsen, RegridFlag.Keep, RegridFlag.Refine


def level(k):
    """The refinement level of a leaf key (patch, level, corner...)."""
    return k[1]


def doublefactorial(m):
    """(2m - 1)!! = 1 * 3 * 5 ... (2m - 1) (refinement.jl:178-180)."""
    return math.prod(range(1, 2 * m, 2))


def check_mode(mode, P):
    if not 1 <= mode <= P:
        raise ValueError(
            f"mode = {mode} is not a Legendre mode an element of order P = {P} can be refined "
            f"by: take P - 1 = {P - 1} for the scheme that truncates v, whose u keeps its top "
            f"mode constant in time, and P = {P} for the plain one. Mode 0 is the element's "
            f"mean, which says nothing about resolution, so the truncated scheme at P = 1 has "
            f"no mode to refine by")


def check_thresholds(refine_tol, coarsen_tol):
    if not coarsen_tol < refine_tol:
        raise ValueError(
            f"coarsen_tol ({coarsen_tol}) must lie below refine_tol ({refine_tol}): the gap "
            f"between them is the dead band that stops blocks from flipping between levels on "
            f"alternate regrids")


def _check_var(var):
    if var not in (1, 2):
        raise ValueError(f"var = {var} is not a variable of this field set, which has 2")


def _scaled_modes(p, modes, weights, pows, post, var, u):
    """post * max over lines and modes of ((weights[q]/A) |c_m|)^pows[q] per element,
    as an array tau[e1, ..., eD, b] (column-major, the reference's layout)."""
    _check_var(var)
    slot = _lib.SLOT_STATE
    if u is not None:
        p.upload(u, _lib.SLOT_ARG)
        slot = _lib.SLOT_ARG
    m = np.asarray(modes, dtype=np.int32)
    w = np.asarray(weights, dtype=np.float64)
    pw = np.asarray(pows, dtype=np.float64)
    tau = np.empty(p.nblocks * p.E ** p.D)
    _lib.check(_lib.lib().twoldg_modes(p._h, slot, var - 1, len(m),
                                       m.ctypes.data_as(ctypes.c_void_p),
                                       w.ctypes.data_as(ctypes.c_void_p),
                                       pw.ctypes.data_as(ctypes.c_void_p), float(post),
                                       tau.ctypes.data_as(ctypes.c_void_p)))
    return tau.reshape((p.E,) * p.D + (p.nblocks,), order="F")


def _check_element(p, el):
    if el is not None and el.P != p.P:
        raise ValueError(f"the field set holds elements of order P = {p.P}, and the element is "
                         f"of order P = {el.P}")


def modal_indicator(p, el=None, *, mode, var=1, u=None):
    """The refinement indicator of every element, tau[e1, ..., eD, b]: with c_m the
    coefficient of L_m of variable `var` along a line of the element and A the
    largest |value| of that variable over the mesh,

        tau = max over the element's lines in every dimension and m = 1 ... mode
              of ((2m - 1)!! |c_m| / A)^(mode/m) / (2 mode - 1)!!

    (refinement.jl:103-155).  It reads the problem's state, or `u` when given (a
    state vector, passed through the problem's argument slot).  `mode` is P - 1
    for the scheme that truncates v and P for the plain one."""
    _check_element(p, el)
    check_mode(mode, p.P)
    M = int(mode)
    modes = list(range(1, M + 1))
    return _scaled_modes(p, modes, [doublefactorial(m) for m in modes],
                         [M / m for m in modes], 1 / doublefactorial(M), var, u)


def mode_amplitude(p, el=None, *, mode, var=1, u=None):
    """The amplitude |c_mode| of the single Legendre mode L_mode of `var`, the
    largest along any line of each element relative to the mesh's largest |value|,
    in modal_indicator's layout (refinement.jl:157-174).  Any mode 0 ... P."""
    _check_element(p, el)
    if not 0 <= mode <= p.P:
        raise ValueError(f"mode = {mode} is not a Legendre mode of an element of order "
                         f"P = {p.P}")
    return _scaled_modes(p, [int(mode)], [1.0], [1.0], 1.0, var, u)


def refine_mark(tau, el, b, k, *, refine_tol, coarsen_tol, maxlevel_cap):
    """The regrid mark of block b with key k from its slice of the indicator
    (refinement.jl:190-253):

        some element above refine_tol, level < cap  ->  (Refine, box)
        some element above coarsen_tol              ->  (Keep, box)
        none, level > 0                             ->  Coarsen
        otherwise                                   ->  Keep

    with `box` the ranges of owned node indices (0-based, end exclusive) of the
    elements above coarsen_tol."""
    check_thresholds(refine_tol, coarsen_tol)
    P1 = el.P + 1
    t = np.asarray(tau)[..., b]
    D = t.ndim
    fire = bool(np.any(t > refine_tol))
    above = np.argwhere(t > coarsen_tol)
    box = None
    if len(above):
        lo, hi = above.min(axis=0), above.max(axis=0)
        box = tuple(range(int(lo[d]) * P1, (int(hi[d]) + 1) * P1) for d in range(D))
    if fire and level(k) < maxlevel_cap:
        return (Refine, box)
    if box is not None:
        return (Keep, box)
    if level(k) > 0:
        return Coarsen
    return Keep


def flag_blocks(forest, fn):
    """[fn(b, k) for every leaf], TreeAMR's flag_blocks."""
    return [fn(b, k) for b, k in enumerate(forest.leaves)]


def refine_flags(p, el=None, *, mode, refine_tol, coarsen_tol, maxlevel_cap, var=1, u=None):
    """refine_mark for every block, from one indicator pass (refinement.jl:255-287)."""
    check_thresholds(refine_tol, coarsen_tol)
    el = el if el is not None else p.element
    tau = modal_indicator(p, el, mode=mode, var=var, u=u)
    return flag_blocks(p.forest, lambda b, k: refine_mark(
        tau, el, b, k, refine_tol=refine_tol, coarsen_tol=coarsen_tol,
        maxlevel_cap=maxlevel_cap))


def refinement_buffer(forest, maxlevel_cap, travel):
    """The regrid buffer in points that covers a feature moving `travel` between
    regrids at the spacing of level maxlevel_cap: ceil(travel/h) + 1, refused above
    N (refinement.jl:289-316)."""
    h = forest.spacing(maxlevel_cap)
    points = math.ceil(travel / h) + 1
    if points > forest.N:
        raise ValueError(
            f"a feature travelling {travel} between regrids needs a {points}-point margin at "
            f"level {maxlevel_cap} (h = {h}), which exceeds the block width N = {forest.N}. "
            f"Recruitment reaches one ring of neighbours, so the travel must stay under one "
            f"finest-level block width, {forest.N * h}: regrid more often.")
    return points


# ---------------------------------------------------------------------------
# The forest of a regrid.
# ---------------------------------------------------------------------------
def _split(flag):
    if isinstance(flag, tuple):
        f, box = flag
        return RegridFlag(f), box
    return RegridFlag(flag), None


def _check_regriddable(forest):
    if forest.geometry is not None or len(forest.patches) != 1:
        raise ValueError(
            "regrid takes the box forests of wave_forest: the glue between the patches of a "
            "multi-patch geometry has no refined counterpart here")


def _unbalanced(forest):
    """Leaves with a face neighbour two or more levels finer."""
    out = set()
    D = forest.D
    for lf in forest.leaves:
        for d in range(D):
            for s in (0, 1):
                nc = forest._neighbour_cell(lf, d, s)
                if nc is None or nc not in forest.internal:
                    continue
                p, lev, *c = nc
                for ch in range(1 << (D - 1)):
                    cc, k = [0, 0, 0], 0
                    for e in range(D):
                        if e == d:
                            cc[e] = 2 * c[e] + (0 if s else 1)
                        else:
                            cc[e] = 2 * c[e] + ((ch >> k) & 1)
                            k += 1
                    if (p, lev + 1) + tuple(cc) in forest.internal:
                        out.add(lf)
    return out


def regrid_forest(forest, flags, buffer=0):
    """The forest after a regrid by per-block `flags` (Refine / Keep / Coarsen, or
    a (flag, box) pair whose box, in owned node indices, recruits the blocks within
    `buffer` points of it to the flagging block's target level), and whether it
    changed.  The input forest is not modified."""
    _check_regriddable(forest)
    D, N = forest.D, forest.N
    if len(flags) != forest.nblocks:
        raise ValueError(f"{len(flags)} flags for {forest.nblocks} blocks")
    if not 0 <= buffer <= N:
        raise ValueError(f"the buffer ({buffer} points) must lie in 0 ... N = {N}: "
                         f"recruitment reaches one ring of neighbours")
    leaves = forest.leaves
    nb = len(leaves)
    want = np.zeros(nb, dtype=np.int64)   # -1 coarsen, 0 keep, 1 refine
    boxes = []
    for b, fl in enumerate(flags):
        f, box = _split(fl)
        if f == Coarsen and level(leaves[b]) == 0:
            f = Keep
        want[b] = int(f)
        if box is not None:
            boxes.append((b, box, level(leaves[b]) + (1 if f == Refine else 0)))

    # Recruitment: every block that meets a dilated box rises to the box's level.
    ext = np.array([[e for e in forest.extent(lf)] for lf in leaves])   # (nb, D, 2)
    roots, plo, phi = forest.patches[0]
    size = np.array([phi[d] - plo[d] for d in range(D)])
    shifts = [np.array(s) for s in itertools.product(
        *[(-1, 0, 1) if forest.periodic[d] else (0,) for d in range(D)])]
    lev = np.array([level(lf) for lf in leaves])
    for b, box, target in boxes:
        (bl, bh) = ext[b, :, 0], ext[b, :, 1]
        hb = (bh[0] - bl[0]) / N
        lo = np.array([bl[d] + (box[d].start - buffer) * hb for d in range(D)])
        hi = np.array([bl[d] + (box[d].stop + buffer) * hb for d in range(D)])
        for sh in shifts:
            a, z = lo + sh * size, hi + sh * size
            hit = np.all((ext[:, :, 0] < z[None, :]) & (ext[:, :, 1] > a[None, :]), axis=1)
            for o in np.nonzero(hit)[0]:
                if o == b:
                    continue
                if lev[o] < target:
                    want[o] = 1
                elif lev[o] == target and want[o] < 0:
                    want[o] = 0

    # Refinement, then coarsening of complete sibling groups that keeps 2:1.
    refined = [lf for b, lf in enumerate(leaves) if want[b] > 0]
    groups = {}
    for b, lf in enumerate(leaves):
        if want[b] < 0:
            p, l, *c = lf
            parent = (p, l - 1) + tuple(x >> 1 for x in c)
            groups.setdefault(parent, []).append(lf)
    groups = {k: v for k, v in groups.items() if len(v) == 1 << D}
    keep_set = set(lf for b, lf in enumerate(leaves) if want[b] == 0)
    keep_set.update(lf for b, lf in enumerate(leaves)
                    if want[b] < 0 and not any(lf in g for g in groups.values()))
    while True:
        new = list(keep_set)
        for (p, l, *c) in refined:
            for ch in range(1 << D):
                cc = [2 * c[d] + ((ch >> d) & 1) if d < D else 0 for d in range(3)]
                new.append((p, l + 1) + tuple(cc))
        for parent in groups:
            new.append(parent)
        cand = forest.with_leaves(new)
        bad = [k for k in _unbalanced(cand) if k in groups]
        if not bad:
            break
        for k in bad:
            keep_set.update(groups.pop(k))
    cand.balance()
    changed = set(cand.leaves) != set(leaves)
    return cand, changed


def transfer_map(src, dst):
    """For every leaf of dst its origin in src, as the C ABI's transfer entries:
    the same leaf (copy), its parent (refine) or its 2^D children (coarsen)."""
    D = dst.D
    out = (_lib.TransferBlock * dst.nblocks)()
    for b, lf in enumerate(dst.leaves):
        e = out[b]
        for i in range(8):
            e.src[i] = -1
        p, l, *c = lf
        if lf in src.index:
            e.kind, e.child, e.src[0] = _lib.XFER_COPY, 0, src.index[lf]
            continue
        parent = (p, l - 1) + tuple(x >> 1 for x in c)
        if l > 0 and parent in src.index:
            e.kind, e.src[0] = _lib.XFER_REFINE, src.index[parent]
            e.child = sum((c[d] & 1) << d for d in range(D))
            continue
        kids = []
        for ch in range(1 << D):
            cc = [2 * c[d] + ((ch >> d) & 1) if d < D else 0 for d in range(3)]
            kids.append((p, l + 1) + tuple(cc))
        if not all(k in src.index for k in kids):
            raise ValueError(f"the leaf {lf} of the regridded forest is more than one level "
                             f"from every leaf of the forest it is regridded from")
        e.kind, e.child = _lib.XFER_COARSEN, 0
        for ch, k in enumerate(kids):
            e.src[ch] = src.index[k]
    return out


def _options(p):
    return dict(truncate=p.truncate, precision=p.precision)


def regrid(p, flags, buffer=0):
    """The regrid of a problem and its state: TreeAMR's regrid!(forest, fs =>
    schedule; flags, buffer) followed by WaveProblem(fs, el) (blast.jl:336-348).
    Returns (q, changed): q a new problem on the regridded forest holding p's
    state moved onto it (refined blocks by the coarse polynomial, coarsened ones
    by the exact L2 projection; both keep sum w u of each variable), or p itself
    when the forest did not change.  As in the reference the state is not
    truncated here: coarsening brings v's top mode back, and the caller truncates
    (truncate_state on a state vector, or q.truncate_state() on the state q holds)."""
    forest, changed = regrid_forest(p.forest, flags, buffer)
    if not changed:
        return p, False
    q = WaveProblem(forest, p.element, **_options(p))
    m = transfer_map(p.forest, forest)
    _lib.check(_lib.lib().twoldg_transfer(q._h, _lib.SLOT_STATE, p._h, _lib.SLOT_STATE,
                                          ctypes.cast(m, ctypes.c_void_p)))
    return q, True


def adapt_to_initial_data(forest, el, *, initial, flags, buffer=0, maxpasses=8,
                          precision="f64"):
    """TreeAMR's adapt_to_initial_data!: fill the initial data at the nodes, flag,
    regrid, and again on the new mesh, until the mesh stops changing or after
    `maxpasses` passes.  `flags(p)` takes the problem holding the data in its state
    and returns the per-block flags.  Returns (forest, passes, converged)."""
    converged = False
    passes = 0
    for passes in range(1, maxpasses + 1):
        p = WaveProblem(forest, el, truncate=False, precision=precision)
        try:
            p.upload(fill_at_nodes(initial, p))
            fl = flags(p)
        finally:
            p.close()
        nxt, changed = regrid_forest(forest, fl, buffer)
        if not changed:
            converged = True
            break
        forest = nxt
    return forest, passes, converged

