"""Forests of blocks and their face descriptors (the host side of the exchange).

Mirrors the reference's forest builders, which sit on TreeAMR:
  wave_forest, refine_regions!, refinement_regions, nested_cubes
      TreeWaveOLDG.jl src/sinewave.jl:114-199
  sevenpatch_forest, refine_cube!
      src/geometry.jl:525-553
  check_layout
      src/layout.jl:31-43

A forest is a list of leaves (patch, level, integer corner at that level) over
patches of root blocks.  Leaves are ordered by patch, by root (first dimension
fastest) and then depth-first in Z order inside each root tree; the state
vector follows that order.  Refinement is face-balanced 2:1 after each level,
which is all the face-only coupling of an element set needs.

`faces()` turns the forest into the C ABI's per-face descriptors: same level
(periodic wrap included), coarser, finer, mirror (reflecting wall), glue
(seven-patch ball) and Dirichlet (the ball's outer sphere).
"""
import copy

import numpy as np

from . import _lib


def check_layout(N, P):
    """Refuse a block size this method cannot run on (layout.jl:31-43)."""
    P1 = P + 1
    if N % P1:
        raise ValueError(
            f"the block size N = {N} is not a whole number of elements of order P = {P}, "
            f"which hold {P1} nodes each: choose N = E*{P1} for some even number of "
            f"elements E per block")
    E = N // P1
    if E % 2:
        raise ValueError(
            f"the block size N = {N} holds E = {E} elements of order P = {P}, and E must be "
            f"even: a block of a refined forest covers half of its parent, and has to cover "
            f"whole elements of it, so N must be a multiple of 2(P+1) = {2 * P1}")


class Forest:
    """Leaves over patches.  `patches[p] = (roots, lo, hi)` per dimension."""

    def __init__(self, D, P, E, patches, periodic, reflecting, geometry=None):
        if E % 2:
            raise ValueError(
                f"E = {E} elements of order P = {P} per block is odd, and E must be even: a "
                f"block of a refined forest covers half of its parent and has to cover whole "
                f"elements of it, so TreeAMR's element centering needs N = {E * (P + 1)} to be "
                f"a multiple of 2(P+1) = {2 * (P + 1)}")
        self.D, self.P, self.E = D, P, E
        self.N = E * (P + 1)
        check_layout(self.N, P)
        self.patches = patches
        self.periodic = tuple(periodic)
        self.reflecting = reflecting
        self.geometry = geometry
        leaves = []
        for p, (roots, lo, hi) in enumerate(patches):
            for r in np.ndindex(*reversed(roots)):
                c = tuple(reversed(r))
                leaves.append((p, 0) + c + (0,) * (3 - D))
        self._set(leaves)

    # -- leaves ------------------------------------------------------------
    def _set(self, leaves):
        self.leaves = sorted(leaves, key=self._order_key)
        self.index = {lf: i for i, lf in enumerate(self.leaves)}
        self.internal = set()
        for (p, lev, *c) in self.leaves:
            c = list(c)
            for l in range(lev - 1, -1, -1):
                c = [x >> 1 for x in c]
                self.internal.add((p, l) + tuple(c))

    def _order_key(self, lf):
        p, lev, *c = lf
        roots = self.patches[p][0]
        root = 0
        s = 1
        for d in range(self.D):
            root += (c[d] >> lev) * s
            s *= roots[d]
        path = []
        for l in range(1, lev + 1):
            bits = 0
            for d in reversed(range(self.D)):
                bits = bits * 2 + ((c[d] >> (lev - l)) & 1)
            path.append(bits)
        return (p, root, tuple(path))

    @property
    def nblocks(self):
        return len(self.leaves)

    def dims(self, p, lev):
        roots = self.patches[p][0]
        return [roots[d] << lev for d in range(self.D)]

    def extent(self, lf):
        p, lev, *c = lf
        roots, lo, hi = self.patches[p]
        out = []
        for d in range(self.D):
            w = (hi[d] - lo[d]) / roots[d] / (1 << lev)
            out.append((lo[d] + w * c[d], lo[d] + w * c[d] + w))
        return out

    def spacings(self):
        """h per block: block extent / N (TreeAMR's spacing)."""
        out = np.empty(self.nblocks)
        for b, (p, lev, *c) in enumerate(self.leaves):
            roots, lo, hi = self.patches[p]
            out[b] = (hi[0] - lo[0]) / roots[0] / (1 << lev) / self.N
        return out

    def minimum_spacing(self):
        return float(self.spacings().min())

    def maxlevel(self):
        return max(lf[1] for lf in self.leaves)

    def spacing(self, level):
        """h at `level` in the first patch (TreeAMR's spacing(forest, level))."""
        roots, lo, hi = self.patches[0]
        return (hi[0] - lo[0]) / roots[0] / (1 << level) / self.N

    def with_leaves(self, leaves):
        """A forest over the same patches holding `leaves`."""
        f = copy.copy(self)
        f._set(list(leaves))
        return f

    # -- refinement ----------------------------------------------------------
    def _refine(self, targets):
        leaves = [lf for lf in self.leaves if lf not in targets]
        for (p, lev, *c) in targets:
            for ch in range(1 << self.D):
                cc = [2 * c[d] + ((ch >> d) & 1) if d < self.D else 0 for d in range(3)]
                leaves.append((p, lev + 1) + tuple(cc))
        self._set(leaves)

    def _neighbour_cell(self, lf, d, s):
        """The cell across face (d, s) at the leaf's level inside its patch, wrapped
        where periodic; None where the face is a patch boundary."""
        p, lev, *c = lf
        c = list(c)
        c[d] += 1 if s else -1
        n = self.dims(p, lev)
        if 0 <= c[d] < n[d]:
            return (p, lev) + tuple(c)
        if self.geometry is None and self.periodic[d]:
            c[d] %= n[d]
            return (p, lev) + tuple(c)
        return None

    def balance(self):
        """Refine until every face neighbour is at most one level finer."""
        while True:
            todo = set()
            for lf in self.leaves:
                for d in range(self.D):
                    for s in (0, 1):
                        nc = self._neighbour_cell(lf, d, s)
                        if nc is None or nc not in self.internal:
                            continue
                        p, lev, *c = nc
                        for ch in range(1 << (self.D - 1)):
                            cc, k = [0, 0, 0], 0
                            for e in range(self.D):
                                if e == d:
                                    cc[e] = 2 * c[e] + (0 if s else 1)
                                else:
                                    cc[e] = 2 * c[e] + ((ch >> k) & 1)
                                    k += 1
                            if (p, lev + 1) + tuple(cc) in self.internal:
                                todo.add(lf)
            if not todo:
                return
            self._refine(todo)

    def refine_regions(self, regions, roots_for_message, patch=None):
        """Region k refines the leaves of level k-1 whose centres lie in it, then
        the forest is balanced (sinewave.jl:141-160)."""
        for k, region in enumerate(regions):
            targets = set()
            for lf in self.leaves:
                if lf[1] != k or (patch is not None and lf[0] != patch):
                    continue
                ext = self.extent(lf)
                if all(region[d][0] < (ext[d][0] + ext[d][1]) / 2 < region[d][1]
                       for d in range(self.D)):
                    targets.add(lf)
            if not targets:
                raise ValueError(
                    f"no block centre of level {k} lies inside the region {region!r} that is "
                    f"to refine it to level {k + 1}; with refined=true the region is the "
                    f"middle half of the box, which holds a block centre only from roots = 3 "
                    f"on (roots = {roots_for_message})")
            self._refine(targets)
            self.balance()

    # -- positions -------------------------------------------------------------
    def node_logical(self, offsets):
        """Logical positions of the owned nodes, (nb, D, N^D): lo + h (B + q(P+1)
        + offsets[i]) with B the block's integer corner in units of h
        (layout.jl:59-76, CODE.md "Positions and quadrature")."""
        D, N, P1 = self.D, self.N, self.P + 1
        idx = np.arange(N)
        qint = (idx // P1) * P1          # q(P+1)
        off = offsets[idx % P1]
        h = self.spacings()
        out = np.empty((self.nblocks, D, N ** D))

        def flat(a, d):
            # a[i_d] broadcast over the block, flattened with the first dimension fastest
            shape = [1] * D
            shape[D - 1 - d] = N
            return np.broadcast_to(a.reshape(shape), (N,) * D).reshape(-1)

        qf, of = [flat(qint, d) for d in range(D)], [flat(off, d) for d in range(D)]
        for b, (p, lev, *c) in enumerate(self.leaves):
            lo = self.patches[p][1]
            for d in range(D):
                out[b, d] = lo[d] + h[b] * ((c[d] * N + qf[d]).astype(np.float64) + of[d])
        return out

    # -- faces -------------------------------------------------------------------
    def faces(self):
        """Per-face descriptors for the C ABI, array of _lib.Face (nb * D * 2)."""
        D = self.D
        out = (_lib.Face * (self.nblocks * D * 2))()
        for b, lf in enumerate(self.leaves):
            for d in range(D):
                for s in (0, 1):
                    F = out[(b * D + d) * 2 + s]
                    for i in range(4):
                        F.nb[i] = -1
                    nc = self._neighbour_cell(lf, d, s)
                    if nc is None:
                        if self.geometry is None:
                            F.kind = _lib.FACE_MIRROR
                        else:
                            self.geometry.glue_face(self, b, d, s, F)
                        continue
                    if nc in self.index:
                        F.kind = _lib.FACE_SAME
                        F.nb[0] = self.index[nc]
                        continue
                    p, lev, *c = nc
                    parent = (p, lev - 1) + tuple(x >> 1 for x in c)
                    if lev > 0 and parent in self.index:
                        F.kind = _lib.FACE_COARSER
                        F.nb[0] = self.index[parent]
                        for e in range(3):
                            F.off[e] = c[e] & 1
                        continue
                    F.kind = _lib.FACE_FINER
                    tdims = [e for e in range(D) if e != d]
                    for ch in range(1 << (D - 1)):
                        cc = [0, 0, 0]
                        cc[d] = 2 * c[d] + (0 if s else 1)
                        for k, e in enumerate(tdims):
                            cc[e] = 2 * c[e] + ((ch >> k) & 1)
                        key = (p, lev + 1) + tuple(cc)
                        if key not in self.index:
                            raise ValueError("the forest is not 2:1 balanced across faces")
                        F.nb[ch] = self.index[key]
        return out


def refinement_regions(refined, D, L):
    """The regions `refined` stands for, one per level (sinewave.jl:163-175)."""

    def isregion(r):
        return (isinstance(r, (tuple, list)) and len(r) == D and
                all(isinstance(i, (tuple, list)) and len(i) == 2 and i[0] < i[1] for i in r))

    if refined is True:
        return [tuple((L / 4, 3 * L / 4) for _ in range(D))]
    if isinstance(refined, tuple) and isregion(refined):
        return [refined]
    if isinstance(refined, list) and all(isregion(r) for r in refined):
        return refined
    raise ValueError(
        f"refined must be false, true, a tuple of {D} intervals (lo, hi) with lo < hi, one per "
        f"dimension, or a vector of such tuples, one per level; got {refined!r}")


def wave_forest(D, P, E, roots=4, L=1.0, reflecting=False, refined=False):
    """A box of roots^D blocks of E^D elements of order P, periodic or reflecting,
    uniform or refined (sinewave.jl:114-135)."""
    forest = Forest(D, P, E, [((roots,) * D, (0.0,) * D, (float(L),) * D)],
                    periodic=(not reflecting,) * D, reflecting=reflecting)
    if refined is False:
        return forest
    forest.refine_regions(refinement_regions(refined, D, L), roots)
    return forest


def nested_cubes(D, w0, levels):
    """Regions of `levels` nested cubes about the centre of a box of side
    2^levels w0 (sinewave.jl:192-199)."""
    if levels < 1:
        raise ValueError(f"a forest has at least one level, got {levels}")
    if not w0 > 0:
        raise ValueError(f"the finest cube's half-width w0 = {w0} must be positive")
    c = 2 ** (levels - 1) * w0
    return [tuple((c - 2 ** (levels - 1 - k) * w0, c + 2 ** (levels - 1 - k) * w0)
                  for _ in range(D)) for k in range(1, levels)]
