"""The seven-patch ball: map, frame, glue and the tables a full metric needs.

Mirrors TreeWaveOLDG.jl src/geometry.jl:
  SevenPatchBall, shell_fraction, blended_shell     geometry.jl:270-381
  geometry_frame (J, S[d][i] = J dxi^d/dx^i)        geometry.jl:402-409
  sevenpatch_glue, radial_glue!, quarter_turns!     geometry.jl:418-455
  sevenpatch_forest, refine_cube!                   geometry.jl:525-553
and the host-side builds of src/evolution.jl that a full metric needs:
  coefficient_set / FrameTerms (sqrt(g), S)         evolution.jl:687-696, 988-998
  flat_blocks                                        evolution.jl:627-634
  block_speeds (hspeed)                              evolution.jl:643-656
  ElementTruncation (weights, Cholesky factors)      evolution.jl:749-899

Patch 0 is the cube [-Lc, Lc]^3 with n roots per side; patches 1..6 are the
shells over the cube's faces (k, s) in the order (1,-1), (1,1), (2,-1), (2,1),
(3,-1), (3,1), each with m roots along its radial dimension k.  Every patch's
logical box is [0, roots]: one logical unit per root block.

The Jacobian is the map's closed-form derivative (the reference differentiates
with ForwardDiff); glued faces are matched by the physical coincidence of the
block faces' corners, which fixes the tangential axis map without assuming any
convention for the glue table's signed permutation.
"""
import math

import numpy as np

from . import _lib
from .forest import Forest, refinement_regions

SHELLS = ((0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1))  # (k, s), 0-based k


def shell_of(p):
    """(k, s) of shell patch p >= 1."""
    return SHELLS[p - 1]


def shell_index(k, s):
    return 1 + 2 * k + (1 if s == 1 else 0)


def _largest_eigenvalue(A):
    """The largest eigenvalue of each symmetric 3x3 matrix of A (n, 3, 3), in
    closed form: with q = tr A / 3 and B = (A - q I) / p, p^2 = |A - q I|_F^2 / 6,
    it is q + 2 p cos(acos(det B / 2) / 3)."""
    q = np.trace(A, axis1=1, axis2=2) / 3
    off = A[:, 0, 1] ** 2 + A[:, 0, 2] ** 2 + A[:, 1, 2] ** 2
    p2 = ((A[:, 0, 0] - q) ** 2 + (A[:, 1, 1] - q) ** 2 + (A[:, 2, 2] - q) ** 2 + 2 * off) / 6
    p = np.sqrt(p2)
    safe = np.where(p > 0, p, 1.0)
    B = (A - q[:, None, None] * np.eye(3)) / safe[:, None, None]
    r = np.clip(np.linalg.det(B) / 2, -1.0, 1.0)
    return np.where(p > 0, q + 2 * p * np.cos(np.arccos(r) / 3), q)


class SevenPatchBall:
    """The ball |x| <= R1 as a cube of half-width Lc and six shells; roots =
    (n, m); grading alpha widens the shells' radial elements outward."""

    def __init__(self, Lc=1.0, R1=3.0, roots=(2, 2), grading=0.0):
        if not Lc > 0:
            raise ValueError(f"the cube's half-width Lc must be positive, got {Lc}")
        if not R1 > math.sqrt(3) * Lc:
            raise ValueError(
                f"the outer radius R1 = {R1} must exceed sqrt(3) Lc = {math.sqrt(3) * Lc}, so "
                f"that the cube's corners lie inside the sphere and every shell maps one to one")
        if not all(r > 0 for r in roots):
            raise ValueError(
                f"every patch of the ball needs a positive number of roots along each "
                f"dimension, got {roots}")
        if not math.isfinite(grading):
            raise ValueError(f"the shells' radial grading must be finite, got {grading}")
        self.Lc, self.R1 = float(Lc), float(R1)
        self.n, self.m = int(roots[0]), int(roots[1])
        self.grading = float(grading)

    def patch_roots(self, p):
        if p == 0:
            return (self.n,) * 3
        k, _ = shell_of(p)
        return tuple(self.m if j == k else self.n for j in range(3))

    def outer_radius(self):
        """The radius of the outer sphere, which a radiative boundary's u/R reads."""
        return self.R1

    def _element_blocks(self, forest):
        """The blocks that take per-element truncation tables: every shell block."""
        return [b for b, lf in enumerate(forest.leaves) if lf[0] != 0]

    # -- the map ------------------------------------------------------------------
    def _rho(self, eta):
        a = self.grading
        if a == 0:
            return eta, np.ones_like(eta)
        ea = math.exp(a) - 1
        return (np.exp(a * eta) - 1) / ea, a * np.exp(a * eta) / ea

    def position(self, p, xi):
        """Physical positions of logical xi (3, n) of patch p."""
        xi = np.asarray(xi, dtype=np.float64)
        if p == 0:
            return (2 * self.Lc / self.n) * xi - self.Lc
        k, s = shell_of(p)
        f = np.empty_like(xi)
        for j in range(3):
            f[j] = s * self.Lc if j == k else self.Lc * (2 * xi[j] / self.n - 1)
        eta = xi[k] / self.m if s == 1 else 1 - xi[k] / self.m
        rho, _ = self._rho(eta)
        nf = np.sqrt(f[0] ** 2 + f[1] ** 2 + f[2] ** 2)
        return (1 - rho) * f + (rho * self.R1 / nf) * f

    def jacobian(self, p, xi):
        """M[i, j] = dx^i / dxi^j at xi (3, n): shape (n, 3, 3)."""
        xi = np.asarray(xi, dtype=np.float64)
        npts = xi.shape[1]
        M = np.zeros((npts, 3, 3))
        a = 2 * self.Lc / self.n
        if p == 0:
            for j in range(3):
                M[:, j, j] = a
            return M
        k, s = shell_of(p)
        f = np.empty_like(xi)
        for j in range(3):
            f[j] = s * self.Lc if j == k else self.Lc * (2 * xi[j] / self.n - 1)
        eta = xi[k] / self.m if s == 1 else 1 - xi[k] / self.m
        rho, drho = self._rho(eta)
        nf = np.sqrt(f[0] ** 2 + f[1] ** 2 + f[2] ** 2)
        # d/dxi_k: rho' (s/m) (R f/|f| - f)
        dk = drho * (s / self.m)
        for i in range(3):
            M[:, i, k] = dk * (self.R1 * f[i] / nf - f[i])
        # d/dxi_j, j != k: (1-rho) a e_j + rho R a/|f| (e_j - f f_j/|f|^2)
        for j in range(3):
            if j == k:
                continue
            for i in range(3):
                proj = (1.0 if i == j else 0.0) - f[i] * f[j] / nf ** 2
                M[:, i, j] = (1 - rho) * a * (1.0 if i == j else 0.0) + rho * self.R1 * a / nf * proj
        return M

    def frame(self, p, xi):
        """(J, S) with J = |det M| (n,) and S[:, d, i] = J (M^-1)[d, i]."""
        M = self.jacobian(p, xi)
        if p == 0:
            a = 2 * self.Lc / self.n
            J = np.full(M.shape[0], a * a * a)
            S = np.zeros_like(M)
            for d in range(3):
                S[:, d, d] = a * a
            return J, S
        # S[d] = J (M^-1)[d, :] = sign(det) (m_{d+1} x m_{d+2}), m_j the columns of M
        m = [M[:, :, j] for j in range(3)]
        S = np.empty_like(M)
        for d in range(3):
            S[:, d, :] = np.cross(m[(d + 1) % 3], m[(d + 2) % 3])
        det = np.einsum("ni,ni->n", m[0], S[:, 0, :])
        sg = np.sign(det)
        return np.abs(det), S * sg[:, None, None]

    def physical(self, forest, x):
        """Physical positions of logical node positions x (nb, 3, nn)."""
        out = np.empty_like(x)
        for b, lf in enumerate(forest.leaves):
            out[b] = self.position(lf[0], x[b])
        return out

    # -- the full metric's tables ---------------------------------------------------
    def metric(self, forest, el):
        """Coefficients (nb, 1 + 9, nn): sqrt(g) then S[d][i] at 1 + 3d + i; flat
        flags; hspeed = min_b h_b / max speed (block_speeds).  Kept for the forest's
        leaf set and P, so that problems of several precisions share one build."""
        key = ("metric", id(forest.leaves), el.P)
        cache = self.__dict__.setdefault("_built", {})
        if key not in cache:
            cache.clear()
            cache[key] = (forest.leaves, self._metric(forest, el))
        return cache[key][1]

    def _metric(self, forest, el):
        x = forest.node_logical(el.offsets)
        nb, _, nn = x.shape
        coef = np.empty((nb, 10, nn))
        speed = np.empty(nb)
        # the frames of many blocks of one patch per call (position and Jacobian
        # take any number of points)
        by_patch = {}
        for b, lf in enumerate(forest.leaves):
            by_patch.setdefault(lf[0], []).append(b)
        chunk = max(1, (1 << 20) // nn)
        for p, blocks in by_patch.items():
            for i in range(0, len(blocks), chunk):
                bs = blocks[i:i + chunk]
                xs = np.moveaxis(x[bs], 1, 0).reshape(3, -1)
                J, S = self.frame(p, xs)
                J = J.reshape(len(bs), nn)
                S = S.reshape(len(bs), nn, 9)
                coef[bs, 0] = J
                coef[bs, 1:] = np.moveaxis(S, 2, 1)
                S3 = S.reshape(-1, 3, 3)
                g = np.einsum("nai,nci->nac", S3, S3) / (J.reshape(-1) ** 2)[:, None, None]
                lam = _largest_eigenvalue(g).reshape(len(bs), nn)
                speed[bs] = np.sqrt(lam.max(axis=1))
        flat = np.zeros(nb, dtype=np.uint8)
        offdiag = [1 + 3 * d + i for d in range(3) for i in range(3) if d != i]
        for b in range(nb):
            c0 = coef[b, :, :1]
            flat[b] = int(np.all(coef[b] == c0) and np.all(coef[b, offdiag, 0] == 0))
        hspeed = float(np.min(forest.spacings() / speed))
        return dict(coef=coef, flat=flat, hspeed=hspeed)

    def truncation_tables(self, forest, el):
        """Per-element sqrt(g)-orthogonal projection tables for the shell blocks:
        slot (nb), modes (n, m), weights (ntables, E^3, n), chol (ntables, E^3, m*m)
        with the upper factor U of G = Psi^T W Psi stored column-major."""
        P, P1, E, N = el.P, el.P1, forest.E, forest.N
        n, m = P1 ** 3, P ** 3
        Lg = el.legendre  # [node, mode]
        k0, k1, k2 = np.meshgrid(np.arange(P1), np.arange(P1), np.arange(P1), indexing="ij")
        q0, q1, q2 = np.meshgrid(np.arange(P), np.arange(P), np.arange(P), indexing="ij")
        # node index k = k0 + P1 k1 + P1^2 k2, mode index q = q0 + P q1 + P^2 q2
        kk = [a.transpose(2, 1, 0).reshape(-1) for a in (k0, k1, k2)]
        qq = [a.transpose(2, 1, 0).reshape(-1) for a in (q0, q1, q2)]
        Psi = (Lg[kk[0][:, None], qq[0][None, :]] * Lg[kk[1][:, None], qq[1][None, :]] *
               Lg[kk[2][:, None], qq[2][None, :]])
        wn = el.weights[kk[0]] * el.weights[kk[1]] * el.weights[kk[2]]
        x = forest.node_logical(el.offsets)
        slot = np.full(forest.nblocks, -1, dtype=np.int32)
        shells = self._element_blocks(forest)
        slot[shells] = np.arange(len(shells), dtype=np.int32)
        ne = E ** 3
        weights = np.empty((len(shells), ne, n))
        chol = np.empty((len(shells), ne, m * m))
        # owned node index of element e's node k
        e0, e1, e2 = np.meshgrid(np.arange(E), np.arange(E), np.arange(E), indexing="ij")
        ee = [a.transpose(2, 1, 0).reshape(-1) for a in (e0, e1, e2)]
        lin = ((ee[0][:, None] * P1 + kk[0][None, :]) +
               N * (ee[1][:, None] * P1 + kk[1][None, :]) +
               N * N * (ee[2][:, None] * P1 + kk[2][None, :]))  # (ne, n)
        for s_, b in enumerate(shells):
            J, _ = self.frame(forest.leaves[b][0], x[b])
            w = wn[None, :] * J[lin]
            weights[s_] = w
            G = np.einsum("kq,ek,kr->eqr", Psi, w, Psi)
            L = np.linalg.cholesky(G)          # G = L L^T, U = L^T
            chol[s_] = L.reshape(ne, m * m)    # U column-major = L row-major
        return dict(slot=slot, ntables=len(shells), modes=np.ascontiguousarray(Psi),
                    weights=np.ascontiguousarray(weights), chol=np.ascontiguousarray(chol))

    def face_positions(self, forest, el):
        """Physical positions of every face node, (nb, 3, 2, 3, N^2): block, face
        dim, side, coordinate, tangential index (the outer ghosts sit there)."""
        N = forest.N
        x = self.physical(forest, forest.node_logical(el.offsets))
        nb = x.shape[0]
        xg = x.reshape(nb, 3, N, N, N)  # [b, c, z, y, x]
        out = np.empty((nb, 3, 2, 3, N * N))
        for d in range(3):
            for s in (0, 1):
                i = N - 1 if s else 0
                if d == 0:
                    pl = xg[:, :, :, :, i]      # [b, c, z, y]: tangential (y, z), y fastest
                elif d == 1:
                    pl = xg[:, :, :, i, :]      # [b, c, z, x]
                else:
                    pl = xg[:, :, i, :, :]      # [b, c, y, x]
                out[:, d, s] = pl.reshape(nb, 3, N * N)
        return out

    # -- glue ------------------------------------------------------------------------
    def neighbour_face(self, p, d, s):
        """The patch face glued to patch p's face (d, side s), or None for the outer
        sphere (geometry.jl:418-455)."""
        if p == 0:
            return shell_index(d, 1 if s else -1), d, 0 if s else 1
        k, sg = shell_of(p)
        if d == k:
            if (sg == 1 and s == 1) or (sg == -1 and s == 0):
                return None  # the sphere
            return 0, k, 1 if sg == 1 else 0
        t = 1 if s else -1
        return shell_index(d, t), k, 1 if sg == 1 else 0

    def _face_corners(self, forest, b, d, s):
        """The physical corners of block b's face (d, s): (0,0), (1,0), (0,1), (1,1) in
        its tangential axes, one position call."""
        p, lev, *c = forest.leaves[b]
        ext = forest.extent(forest.leaves[b])
        tdims = [e for e in range(3) if e != d]
        xi = np.zeros((3, 4))
        xi[d] = ext[d][1] if s else ext[d][0]
        xi[tdims[0]] = [ext[tdims[0]][0], ext[tdims[0]][1]] * 2
        xi[tdims[1]] = [ext[tdims[1]][0]] * 2 + [ext[tdims[1]][1]] * 2
        return self.position(p, xi).T

    def _centre_key(self, corners):
        return tuple(np.round(corners.mean(axis=0) / (1e-6 * self.R1)).astype(np.int64))

    def _face_blocks(self, forest, q, lev, dq, sq):
        """The leaves of patch q at level lev on that patch's face (dq, sq), from an
        index built once per forest leaf set."""
        return self._face_index(forest)[0].get((q, lev, dq, sq), [])

    def _face_index(self, forest):
        """Per (patch, level, face) the leaves on that patch face, and per such face
        its corners keyed by their rounded centre; built once per leaf set."""
        idx = getattr(forest, "_face_index_cache", None)
        if idx is None or idx[0] is not forest.leaves:
            table, centres = {}, {}
            for cb, lf in enumerate(forest.leaves):
                p, l, *c = lf
                n = forest.dims(p, l)
                for e in range(3):
                    for side in (0, 1):
                        if c[e] != (n[e] - 1 if side else 0):
                            continue
                        table.setdefault((p, l, e, side), []).append(cb)
                        corners = self._face_corners(forest, cb, e, side)
                        centres[(p, l, e, side) + self._centre_key(corners)] = (cb, corners)
            idx = (forest.leaves, table, centres)
            forest._face_index_cache = idx
        return idx[1], idx[2]

    def glue_face(self, forest, b, d, s, F):
        p = forest.leaves[b][0]
        nf = self.neighbour_face(p, d, s)
        if nf is None:
            F.kind = _lib.FACE_DIRICHLET
            return
        q, dq, sq = nf
        mine = self._face_corners(forest, b, d, s)
        tol = 1e-9 * self.R1
        lev = forest.leaves[b][1]
        table, centres = self._face_index(forest)
        hit = centres.get((q, lev, dq, sq) + self._centre_key(mine))
        cands = [hit] if hit is not None else []
        cands += [(cb, None) for cb in table.get((q, lev, dq, sq), [])]
        for cb, theirs in cands:
            if theirs is None:
                theirs = self._face_corners(forest, cb, dq, sq)
            match = []
            for c in mine:
                dist = np.linalg.norm(theirs - c, axis=1)
                j = int(np.argmin(dist))
                if dist[j] > tol:
                    break
                match.append(j)
            if len(match) != 4:
                continue
            # their corner index j = a0 + 2 a1; ours are (0,0), (1,0), (0,1), (1,1)
            bits = [(j & 1, j >> 1) for j in match]
            F.kind = _lib.FACE_GLUE
            F.nb[0] = cb
            F.nface = 2 * dq + sq
            for jax in range(2):
                if bits[1][jax] != bits[0][jax]:      # moves along our axis 0
                    F.map[jax] = 1 if bits[1][jax] > bits[0][jax] else -1
                else:                                  # along our axis 1
                    F.map[jax] = 2 if bits[2][jax] > bits[0][jax] else -2
            return
        raise ValueError(f"no block glued to block {b}'s face ({d}, {s})")


def sevenpatch_forest(P, E, roots=(2, 2), Lc=1.0, R1=3.0, refined=False, grading=0.0):
    """The seven-patch ball as a forest and its map (geometry.jl:525-538)."""
    geo = SevenPatchBall(Lc, R1, roots, grading)
    patches = []
    for p in range(7):
        r = geo.patch_roots(p)
        patches.append((r, (0.0,) * 3, tuple(float(x) for x in r)))
    forest = Forest(3, P, E, patches, periodic=(False,) * 3, reflecting=False, geometry=geo)
    if refined is not False:
        forest.refine_regions(refinement_regions(refined, 3, geo.n), geo.n, patch=0)
        # TreeAMR keeps a glued face conforming: a refined cube block on the
        # cube's boundary would refine the shell across it
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
