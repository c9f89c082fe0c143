"""CPU tests of the host side: the C ABI library loads and exports every
declared symbol, the tables are the oracle's bit for bit, the forests and the
exchange plan reproduce the oracle's, and refusals carry the reference's
reasons.  No GPU compute is called."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
import twoldg as T
from twoldg import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "twoldg.h")).read()
    declared = set(re.findall(r"^\s*(?:int|const char\*)\s+(twoldg_\w+)\(", header, re.M))
    assert declared == set(_lib.exported_symbols())
    L = ctypes.CDLL(_lib.LIBPATH)
    for name in declared:
        assert hasattr(L, name), name
    assert _lib.lib().twoldg_abi_version() == _lib.ABI_VERSION


@pytest.mark.parametrize("P", range(1, 9))
def test_tables_equal_oracle_bitwise(P):
    """512-bit tables rounded once (product) = binary128 tables rounded once
    (oracle), element.jl:122-137's precision rule."""
    el = T.Element(P)
    ref = O.element(P)
    for name, mine in [("nodes", el.nodes), ("weights", el.weights), ("offsets", el.offsets),
                       ("stencil", el.stencil), ("trunc", el.truncate),
                       ("topmode", el.topmode), ("topdual", el.topdual)]:
        assert np.array_equal(mine, ref[name]), name


def _plan_vs_oracle(D, P, E, roots, reflecting=False, regions=()):
    f = T.wave_forest(D, P, E, roots=roots, reflecting=reflecting,
                      refined=list(regions) if regions else False)
    o = O.Problem.brick(D, P, E, roots=roots, reflecting=reflecting, regions=regions)
    ob = o.blocks()
    assert [(lf[0], lf[1], tuple(lf[2:2 + D])) for lf in f.leaves] == [b[:3] for b in ob]
    assert np.array_equal(f.spacings(), np.array([b[3] for b in ob]))
    assert np.array_equal(f.node_logical(T.Element(P).offsets), o.coords())
    # the C ABI's host-side positions, bitwise
    from twoldg.problem import node_coordinates
    assert np.array_equal(node_coordinates(f), o.coords())
    y = np.random.default_rng(1).standard_normal(o.size())
    ug, gg = o.ghosts(y)
    g = o.gradients(y).reshape(D, f.nblocks, -1)
    plan = T.exchange_plan(f)
    nn, Nt = (E * (P + 1)) ** D, (E * (P + 1)) ** (D - 1)
    rp, cb, cl = plan["rowptr"], plan["cb"].astype(np.int64), plan["cl"]
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    mu = np.zeros(len(rp) - 1)
    np.add.at(mu, rows, plan["vu"] * y[cb * 2 * nn + cl])
    mg = np.zeros(len(rp) - 1)
    np.add.at(mg, rows, plan["vg"] * g[(rows // (2 * Nt)) % D, cb, cl])
    return np.max(np.abs(mu - ug)), np.max(np.abs(mg - gg))


@pytest.mark.parametrize("case", [
    (2, 4, 2, 8, False, ()),
    (3, 3, 2, 4, False, ()),
    (2, 2, 2, 4, True, ()),
])
def test_exchange_plan_uniform_is_oracle_bitwise(case):
    D, P, E, roots, refl, reg = case
    eu, eg = _plan_vs_oracle(D, P, E, roots, refl, reg)
    assert eu == 0 and eg == 0


@pytest.mark.parametrize("case", [
    (1, 3, 2, 4, False, [((0.25, 0.75),)]),
    (2, 3, 2, 4, False, [((0.25, 0.75),) * 2]),
    (3, 2, 2, 4, False, [((0.25, 0.75),) * 3]),
    (3, 3, 2, 8, True, O.nested_cubes(3, 1 / 8, 3)),
])
def test_exchange_plan_refined_matches_oracle(case):
    """Coarse-fine ghosts: tabulated I and R (product) against Lagrange
    polynomials evaluated on the fly (oracle), to rounding."""
    D, P, E, roots, refl, reg = case
    eu, eg = _plan_vs_oracle(D, P, E, roots, refl, reg)
    assert eu < 1e-13 and eg < 1e-10


def test_nested_cubes_forest_counts():
    """C3: 448 + 448 + 512 leaves (nested_tests.jl:54-80)."""
    f = T.wave_forest(3, 3, 2, roots=8, reflecting=True, refined=T.nested_cubes(3, 1 / 8, 3))
    levels = [lf[1] for lf in f.leaves]
    assert [levels.count(l) for l in range(3)] == [448, 448, 512]
    assert f.minimum_spacing() == pytest.approx(1 / 256)


def test_layout_refusals():
    with pytest.raises(ValueError, match="E must be even"):
        T.wave_forest(2, 3, 3)
    with pytest.raises(ValueError, match="at least 1"):
        T.Element(0)
    with pytest.raises(ValueError, match="no block centre"):
        T.wave_forest(2, 2, 2, roots=2, refined=True)
    with pytest.raises(ValueError, match="refined must be"):
        T.wave_forest(2, 2, 2, roots=4, refined="yes")


@pytest.mark.parametrize("roots", [(2, 2), (4, 4)])
def test_ball_geometry_and_glue_match_oracle(roots):
    """Seven-patch ball: positions bitwise, the closed-form frame against the
    oracle's dual-number Jacobian, and every glued ghost (identity and quarter
    turns, geometry.jl:418-455) equal to the oracle's physical-coincidence
    source, through the exchange plan."""
    from twoldg.geometry import sevenpatch_forest
    P, E = 3, 2
    f, geo = sevenpatch_forest(P, E, roots=roots)
    o = O.Problem.ball(P, E, roots=roots)
    assert [(lf[0], lf[1], tuple(lf[2:5])) for lf in f.leaves] == [b[:3] for b in o.blocks()]
    el = T.Element(P)
    assert np.array_equal(geo.physical(f, f.node_logical(el.offsets)), o.coords())
    m = geo.metric(f, el)
    co, flo = o.coef()
    assert np.max(np.abs(m["coef"] - co)) <= 1e-14 * np.max(np.abs(co))
    assert np.array_equal(m["flat"].astype(bool), flo) and m["flat"].sum() == roots[0] ** 3
    assert m["hspeed"] == pytest.approx(o.timestep(1.0) * (P + 1), rel=1e-14)
    plan = T.exchange_plan(f)
    y = np.random.default_rng(1).standard_normal(o.size())
    ug, gg = o.ghosts(y, 0.0)
    nn, Nt = (E * (P + 1)) ** 3, (E * (P + 1)) ** 2
    rp, cb, cl = plan["rowptr"], plan["cb"].astype(np.int64), plan["cl"]
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    has = np.diff(rp) > 0
    assert (~has).sum() == 6 * roots[0] ** 2 * Nt      # the sphere's ghosts take data
    mu = np.zeros(len(rp) - 1)
    np.add.at(mu, rows, plan["vu"] * y[cb * 2 * nn + cl])
    assert np.array_equal(mu[has], ug[has])
    g = o.gradients(y, 0.0).reshape(f.nblocks, 3, nn)
    ggm = gg.reshape(-1, 3, Nt)
    for i in range(3):
        mg = np.zeros(len(rp) - 1)
        np.add.at(mg, rows, plan["vu"] * g[cb, i, cl])
        hf = has.reshape(-1, Nt)
        assert np.array_equal(mg.reshape(-1, Nt)[hf], ggm[:, i, :][hf])


def test_ball_refusals():
    from twoldg.geometry import SevenPatchBall, sevenpatch_forest
    with pytest.raises(ValueError, match="must exceed"):
        SevenPatchBall(1.0, 1.5, (2, 2))
    with pytest.raises(ValueError, match="reaches the cube's faces"):
        sevenpatch_forest(2, 2, roots=(2, 2), refined=((0.0, 2.0),) * 3)
    f, geo = sevenpatch_forest(2, 2, roots=(4, 2), refined=True)
    assert sum(1 for lf in f.leaves if lf[1] == 1) == 64


GOLDEN = os.path.join(ROOT, "tests", "golden")


def test_oracle_reproduces_golden_fixtures():
    """The committed fixtures (tests/golden/gen_golden.py) are the oracle's
    output; the oracle is deterministic, so it reproduces them exactly."""
    import json
    f = np.load(os.path.join(GOLDEN, "rhs_2d_p4.npz"))
    pr = O.Problem.brick(2, 4, 2, roots=2)
    assert np.array_equal(pr.rhs(f["y"]), f["du"])
    f = np.load(os.path.join(GOLDEN, "rhs_ball_p1.npz"))
    pb = O.Problem.ball(1, 2, roots=(2, 2))
    assert np.array_equal(pb.rhs(f["y"], 0.25), f["du"])
    assert np.array_equal(pb.truncate(f["y"]), f["ytrunc"])
    res = json.load(open(os.path.join(GOLDEN, "results.json")))
    r = O.wave_errors(2, 4, 2, roots=8)
    for k, v in res["wave_d2_p4_r8"].items():
        assert r[k] == pytest.approx(v, rel=1e-14)


def _batch_desc(forest, precision=_lib.F64, keep=None):
    sp = forest.spacings()
    faces = forest.faces()
    d = _lib.Desc()
    d.abi_version = _lib.ABI_VERSION
    d.D, d.P, d.E = forest.D, forest.P, forest.E
    d.precision = precision
    d.truncate = 1
    d.nblocks = forest.nblocks
    d.spacing = sp.ctypes.data_as(ctypes.c_void_p)
    d.faces = ctypes.cast(faces, ctypes.c_void_p)
    if keep is not None:
        keep += [sp, faces, d]
    return d


def test_batch_create_refusals():
    """twoldg_batch_create checks every description before it reads any (no GPU
    call happens on a refusal): a missing one, an unknown precision, no blocks, a
    neighbour that does not face back."""
    keep = []
    f = T.wave_forest(2, 2, 2, roots=2)
    good = _batch_desc(f, keep=keep)
    bad_prec = _batch_desc(f, precision=7, keep=keep)
    no_blocks = _batch_desc(f, keep=keep)
    no_blocks.nblocks = 0
    # block 0's x-high neighbour made block 2, whose x-low neighbour is block 3:
    # block 0's face could not reach block 2's ghost slots
    one_way = _batch_desc(f, keep=keep)
    faces = f.faces()
    assert faces[0 * 4 + 1].nb[0] == 1 and faces[2 * 4 + 0].nb[0] == 3
    faces[0 * 4 + 1].nb[0] = 2
    one_way.faces = ctypes.cast(faces, ctypes.c_void_p)
    keep.append(faces)
    dt = np.array([0.01, 0.01])
    out = ctypes.c_void_p()
    L = _lib.lib()
    for second, msg in [(None, "problem 1: no description"),
                        (bad_prec, "problem 1: unknown precision 7"),
                        (no_blocks, "problem 1: a problem needs at least one block"),
                        (one_way, "problem 1: block 0 face 1: its neighbour does not face back")]:
        arr = (ctypes.c_void_p * 2)()
        arr[0] = ctypes.cast(ctypes.pointer(good), ctypes.c_void_p)
        arr[1] = ctypes.cast(ctypes.pointer(second), ctypes.c_void_p) if second else None
        rc = L.twoldg_batch_create(arr, 2, dt.ctypes.data_as(ctypes.c_void_p), ctypes.byref(out))
        assert rc == -1
        assert L.twoldg_last_error().decode().startswith(msg), L.twoldg_last_error()
    arr = (ctypes.c_void_p * 1)()
    arr[0] = None
    assert L.twoldg_batch_create(arr, 1, dt.ctypes.data_as(ctypes.c_void_p),
                                 ctypes.byref(out)) == -1


def test_convergence_rate_closed_form():
    """errs = C h^p gives the rate p (evolution.jl:1534-1540); a perturbed
    sequence gives the least-squares slope."""
    hs = [1 / 4, 1 / 8, 1 / 16, 1 / 32]
    assert T.convergence_rate(hs, [3.0 * h ** 4 for h in hs]) == pytest.approx(4, abs=1e-12)
    errs = [2.0 * h ** 3 * (1.1 if i % 2 else 1.0) for i, h in enumerate(hs)]
    x, y = np.log(hs), np.log(errs)
    assert T.convergence_rate(hs, errs) == pytest.approx(np.polyfit(x, y, 1)[0], rel=1e-12)


@pytest.mark.parametrize("roots", [(2, 1), (4, 2)])
def test_annulus_forest_metric_tables_match_oracle(roots):
    """The annulus on the host (leaf order, faces: identity glue in theta, walls in
    r; sqrt(g), A_d and hspeed; the per-line truncation pairs built in 40-digit
    arithmetic) equals the oracle's binary128 build bit for bit, and the exact
    mode equals the oracle's to the Bessel functions' rounding."""
    P = 3
    forest, geo = T.annulus_forest(P, 2, roots=roots)
    ref = O.Problem.annulus(P, 2, roots=roots)
    assert [(lf[0], lf[2], lf[3]) for lf in forest.leaves] == \
        [(b[0], b[2][0], b[2][1]) for b in ref.blocks()]
    kind, nb = ref.faces()
    F = forest.faces()
    for b in range(forest.nblocks):
        for d in range(2):
            for s in range(2):
                f = F[(b * 2 + d) * 2 + s]
                assert f.kind == kind[b, d, s]
                if f.kind == _lib.FACE_SAME:
                    assert f.nb[0] == nb[b, d, s]
    el = T.Element(P)
    m = geo.metric(forest, el)
    c, lm, ld = ref.dcoef()
    assert np.array_equal(m["dcoef"].reshape(c.shape), c)
    assert m["hspeed"] == O.lib().orc_hspeed(ref.h)
    tt = geo.truncation_tables(forest, el)
    assert np.array_equal(tt["lmode"], lm) and np.array_equal(tt["ldual"], ld)
    k = T.annulus_wavenumber(2, 1, 1.0, 2.0)
    x = T.problem.fill_at_nodes_forest(forest, T.annulus_exact(geo, forest, 2, k, 0.3), el)
    assert np.max(np.abs(x - ref.fill_exact(4, 0.3, (2, k)))) < 1e-14


def test_annulus_refusals():
    with pytest.raises(ValueError, match="0 < r0 < r1"):
        T.annulus_forest(3, 2, r0=2.0, r1=1.0)
    with pytest.raises(ValueError, match="E must be even"):
        T.annulus_forest(3, 3)
    with pytest.raises(ValueError, match="m >= 0 and n >= 1"):
        T.annulus_wavenumber(1, 0, 1.0, 2.0)
