"""The entry point that allocates nothing: twoldg_norms_noalloc against
twoldg_norms through the C ABI.

Both run the same kernels in the same order, one with the problem's partial-sum
buffers and one with the library's fixed scratch, so the results are equal bit
for bit.  Two calls in a row on the same GPU, the second on a smaller problem
(the scratch keeps the first call's partials: nothing may depend on them), and
a problem beyond the scratch's limit refused on the host.
"""
import ctypes

import numpy as np
import pytest

import twoldg as T
from twoldg import _lib

pytestmark = pytest.mark.gpu


def _call(fn, p, a, b, mask, pn):
    lp = ctypes.c_double()
    li = ctypes.c_double()
    rc = fn(p._h, a, b, mask, pn, ctypes.byref(lp), ctypes.byref(li))
    return rc, lp.value, li.value


def test_norms_noalloc_equals_norms_on_two_calls():
    L = _lib.lib()
    rng = np.random.default_rng(41)
    big = T.WaveProblem(T.wave_forest(3, 3, 2, roots=4), T.Element(3))                # 64 blocks
    small = T.WaveProblem(T.wave_forest(2, 2, 2, roots=2, reflecting=True), T.Element(2),
                          precision="f32c")                                           # 4 blocks
    for p in (big, small):
        u = rng.standard_normal(p.state_len)
        w = rng.standard_normal(p.state_len)
        p.upload(u, _lib.SLOT_ARG)
        p.upload(w, _lib.SLOT_AUX)
        for a, b, mask, pn in [(_lib.SLOT_ARG, _lib.SLOT_AUX, 3, 2), (_lib.SLOT_ARG, -1, 1, 2),
                               (_lib.SLOT_AUX, _lib.SLOT_ARG, 2, 3), (_lib.SLOT_ARG, -1, 3, 1)]:
            rc0, lp0, li0 = _call(L.twoldg_norms, p, a, b, mask, pn)
            rc1, lp1, li1 = _call(L.twoldg_norms_noalloc, p, a, b, mask, pn)
            assert rc0 == 0 and rc1 == 0, L.twoldg_last_error()
            assert (lp1, li1) == (lp0, li0), (p.nblocks, a, b, mask, pn)
            assert lp1 > 0 and li1 > 0
    # the same norm through the Python front end, on the small problem last used
    assert _call(L.twoldg_norms_noalloc, small, _lib.SLOT_ARG, -1, 3, 2)[1] == \
        pytest.approx(T.quadrature_norm(small, small.download(_lib.SLOT_ARG)), rel=1e-15)
    big.close()
    small.close()


def test_norms_noalloc_refuses_beyond_its_scratch():
    """131073 blocks (one more than the scratch holds) are refused before any
    launch; twoldg_norms takes them."""
    L = _lib.lib()
    p = T.WaveProblem(T.wave_forest(1, 1, 2, roots=131073), T.Element(1))
    p.upload(np.ones(p.state_len), _lib.SLOT_ARG)
    rc, _, _ = _call(L.twoldg_norms_noalloc, p, _lib.SLOT_ARG, -1, 3, 2)
    assert rc == -1
    assert "at most 131072 blocks" in L.twoldg_last_error().decode()
    rc, lp, li = _call(L.twoldg_norms, p, _lib.SLOT_ARG, -1, 3, 2)
    assert rc == 0 and lp == pytest.approx(1.0, rel=1e-14) and li == 1.0
    p.close()
