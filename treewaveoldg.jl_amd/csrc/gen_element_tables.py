"""Generate the reference-element tables of the OLDG method in 512-bit precision.

Every coefficient is built in mpmath at 512 bits and rounded to binary64 (and
binary32) once, as the reference does with BigFloat
(TreeWaveOLDG.jl src/element.jl:95-170).  The output is

* ``element_tables.h``  -- C++ constexpr tables read by the HIP library;
* ``../twoldg/_element_tables.py`` -- the same numbers for the Python host.

Tables, for P = 1..8 (``P1 = P + 1`` nodes per element):

nodes, weights, offsets      GLL nodes on [-1,1], weights, (P+1)(1+xi)/2
stencil[P1][P1+2]            wide derivative W in units of 1/h
                             (element.jl:139-151)
trunc[P1][P1]                projector I - L_P (x) dual_P (element.jl:153-155)
topmode[P1], topdual[P1]     the projector's rank-one pair
legendre[P1][P1]             L_k(xi_i), row i = node, column k = mode
interp[2*P1][P1]             coarse element's degree-P Lagrange basis at the
                             2(P+1) fine nodes covering it: fine sub-element
                             f in {0,1}, node j sits at (xi_j + 2f - 1)/2
restrict[P1][2*P1]           R = M_c^-1 I^T M_f, the quadrature adjoint of
                             interp (the coarse ghost of a coarse-fine face)
duals[P1][P1]                row k: the dual of Legendre mode k (its coefficient
                             from nodal values); row P is topdual

Run:  python gen_element_tables.py   (writes both files next to this one)
"""
import os
import mpmath as mp

PMAX = 8
mp.mp.prec = 512


def legendre_all(P, x):
    """L_0..L_P at x by the three-term recurrence (element.jl:72-80)."""
    L = [mp.mpf(1)]
    if P >= 1:
        L.append(x)
    for k in range(2, P + 1):
        L.append(((2 * k - 1) * x * L[k - 1] - (k - 1) * L[k - 2]) / k)
    return L


def gll(P):
    """GLL nodes and weights: Newton on (1-x^2)L_P' from Chebyshev points,
    symmetrized, endpoints exactly +-1 (element.jl:95-114)."""
    x = [-mp.cos(mp.pi * i / P) for i in range(P + 1)]
    tol = mp.ldexp(mp.mpf(1), 8 - mp.mp.prec)
    for _ in range(200):
        delta = []
        for xi in x:
            L = legendre_all(P, xi)
            delta.append((xi * L[P] - L[P - 1]) / ((P + 1) * L[P]))
        x = [xi - di for xi, di in zip(x, delta)]
        if max(abs(d) for d in delta) <= tol:
            break
    x = [(x[i] - x[P - i]) / 2 for i in range(P + 1)]
    x[0], x[P] = mp.mpf(-1), mp.mpf(1)
    w = [2 / (P * (P + 1) * legendre_all(P, xi)[P] ** 2) for xi in x]
    return x, w


def lagrange(nodes, i, x):
    r = mp.mpf(1)
    for j, xj in enumerate(nodes):
        if j != i:
            r *= (x - xj) / (nodes[i] - xj)
    return r


def element(P):
    P1 = P + 1
    xi, w = gll(P)
    Lall = [legendre_all(P, x) for x in xi]
    LP = [L[P] for L in Lall]
    d = [[mp.mpf(0)] * P1 for _ in range(P1)]
    for i in range(P1):
        for j in range(P1):
            if i != j:
                d[i][j] = LP[i] / (LP[j] * (xi[i] - xi[j]))
    d[0][0] = -mp.mpf(P) * (P + 1) / 4
    d[P][P] = mp.mpf(P) * (P + 1) / 4
    W = [[mp.mpf(0)] * (P1 + 2) for _ in range(P1)]
    for i in range(P1):
        for j in range(P1):
            W[i][j + 1] = 2 * d[i][j] / (P + 1)
    face = mp.mpf(P) / 2
    W[0][1] += face
    W[0][0] -= face
    W[P][P1] -= face
    W[P][P1 + 1] += face
    gamma = mp.fsum(w[i] * LP[i] ** 2 for i in range(P1))
    trunc = [[(1 if i == j else 0) - LP[i] * w[j] * LP[j] / gamma for j in range(P1)]
             for i in range(P1)]
    topdual = [w[j] * LP[j] / gamma for j in range(P1)]
    # duals[k][j]: the coefficient of L_k in the element's polynomial through
    # values u is sum_j duals[k][j] u_j (discrete orthogonality under the
    # quadrature); the last row is topdual itself (element.jl:154-159)
    duals = [[w[j] * Lall[j][k] / mp.fsum(w[i] * Lall[i][k] ** 2 for i in range(P1))
              for j in range(P1)] for k in range(P)] + [topdual]
    offsets = [(P + 1) * (1 + x) / 2 for x in xi]
    offsets[0], offsets[P] = mp.mpf(0), mp.mpf(P1)
    interp = []
    for f in range(2):
        for j in range(P1):
            xc = (xi[j] + 2 * f - 1) / 2
            interp.append([lagrange(xi, i, xc) for i in range(P1)])
    restrict = [[interp[f * P1 + j][i] * w[j] / (2 * w[i]) for f in range(2) for j in range(P1)]
                for i in range(P1)]
    return dict(nodes=xi, weights=w, offsets=offsets, stencil=W, trunc=trunc,
                topmode=LP, topdual=topdual, legendre=[[Lall[i][k] for k in range(P1)]
                                                       for i in range(P1)],
                interp=interp, restrict=restrict, duals=duals)


def f64hex(x):
    return float(mp.mpf(x)).hex()


def flat(v):
    if isinstance(v, list) and v and isinstance(v[0], list):
        return [y for row in v for y in row]
    return list(v)


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    els = {P: element(P) for P in range(1, PMAX + 1)}
    names = ["nodes", "weights", "offsets", "stencil", "trunc", "topmode", "topdual",
             "legendre", "interp", "restrict", "duals"]
    # C++ header: one struct per P, padded to the largest size.
    h = ["// Generated by gen_element_tables.py: OLDG reference-element tables,",
         "// built at 512 bits and rounded once to binary64.  Do not edit.",
         "#pragma once", "", "namespace twoldg {", "",
         f"constexpr int kMaxP = {PMAX};", "",
         "struct ElementTables {",
         "  int P;",
         "  double nodes[9], weights[9], offsets[9];",
         "  double stencil[9][11];   // row i, column j+1 for j = -1..P+1",
         "  double trunc[9][9];",
         "  double topmode[9], topdual[9];",
         "  double legendre[9][9];   // [node][mode]",
         "  double interp[18][9];    // [fine node][coarse node]",
         "  double restrict_[9][18]; // [coarse node][fine node]",
         "  double duals[9][9];      // [mode][node]",
         "};", "",
         f"inline constexpr ElementTables kElementTables[{PMAX}] = {{"]
    for P in range(1, PMAX + 1):
        e = els[P]
        P1 = P + 1

        def vec(v):
            return "{" + ", ".join(f64hex(x) for x in v) + "}"

        def mat(m):
            return "{" + ", ".join(vec(r) for r in m) + "}"
        h.append("  {" + f"{P}, " + ", ".join([
            vec(e["nodes"]), vec(e["weights"]), vec(e["offsets"]), mat(e["stencil"]),
            mat(e["trunc"]), vec(e["topmode"]), vec(e["topdual"]), mat(e["legendre"]),
            mat(e["interp"]), mat(e["restrict"]), mat(e["duals"])]) + "},")
    h += ["};", "", "}  // namespace twoldg", ""]
    with open(os.path.join(here, "element_tables.h"), "w") as fh:
        fh.write("\n".join(h))
    # Python module: hex strings per table, row-major.
    py = ['"""Generated by csrc/gen_element_tables.py: OLDG reference-element tables',
          'built at 512 bits and rounded once to binary64 (hex literals).  Do not edit."""',
          "", "TABLES = {"]
    for P in range(1, PMAX + 1):
        e = els[P]
        py.append(f"    {P}: {{")
        for n in names:
            v = e[n]
            shape = (len(v), len(v[0])) if isinstance(v[0], list) else (len(v),)
            py.append(f"        {n!r}: ({shape!r}, {[f64hex(x) for x in flat(v)]!r}),")
        py.append("    },")
    py += ["}", ""]
    with open(os.path.join(here, "..", "twoldg", "_element_tables.py"), "w") as fh:
        fh.write("\n".join(py))


if __name__ == "__main__":
    main()
