"""The OLDG reference element (TreeWaveOLDG.jl src/element.jl:58-170).

The tables are the 512-bit ones of ``csrc/gen_element_tables.py`` rounded once
to binary64 (and from there to binary32 where the FP32 form asks), the
reference's precision rule (element.jl:21-22, CODE.md "Precision").
"""
import numpy as np

from ._element_tables import TABLES

PMAX = 8


def _table(P, name):
    shape, hexes = TABLES[P][name]
    return np.array([float.fromhex(x) for x in hexes], dtype=np.float64).reshape(shape)


class Element:
    """Element(P): nodes, weights, offsets, stencil ((P+1) x (P+3), per h),
    truncate, topmode, topdual, legendre (L_k at the nodes), and the
    coarse-fine operators interp (I) and restrict (R = M_c^-1 I^T M_f).

    Refuses P < 1 as element.jl:96 does."""

    def __init__(self, P):
        if not (isinstance(P, (int, np.integer)) and P >= 1):
            raise ValueError(f"the element order P must be at least 1, got {P}")
        if P > PMAX:
            raise ValueError(f"the element order P must be at most {PMAX}, got {P}")
        self.P = int(P)
        self.P1 = self.P + 1
        self.nodes = _table(P, "nodes")
        self.weights = _table(P, "weights")
        self.offsets = _table(P, "offsets")
        self.stencil = _table(P, "stencil")
        self.truncate = _table(P, "trunc")
        self.topmode = _table(P, "topmode")
        self.topdual = _table(P, "topdual")
        self.legendre = _table(P, "legendre")
        self.interp = _table(P, "interp")
        self.restrict = _table(P, "restrict")

    def __repr__(self):
        return f"Element(P={self.P})"


def order(el):
    return el.P
