"""GPU parity of the complex128 path (ztgevc3) and the complex64 interface (ctgevc3)."""
import numpy as np
import pytest

from tests.helpers import check_cplx_parity, run_cplx
from xtgevc3_amd import pencils as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 2, 9, 31, 32, 33, 64, 100, 161])
@pytest.mark.parametrize("gen", ["ref", "rand", "qz"])
def test_complex_both_back(gen, n):
    if gen == "ref":
        S, P, al, be = G.reference_style_complex(n, 40 + n)
    elif gen == "rand":
        S, P, al, be = G.complex_triangular(n, 50 + n)
    else:
        if n < 2:
            pytest.skip("qz needs n >= 2")
        S, P, Qq, Zq, al, be = G.qz_pencil(n, 60 + n, complex_=True)
    Q = G.random_unitary(n, 3) if gen != "qz" else Qq
    Z = G.random_unitary(n, 4) if gen != "qz" else Zq
    check_cplx_parity(S, P, al, be, "B", "B", None, Q, Z, cond=n <= 100)


@pytest.mark.parametrize("howmny", ["A", "S"])
@pytest.mark.parametrize("side", ["R", "L", "B"])
def test_complex_selected(side, howmny):
    n = 120
    S, P, al, be = G.complex_triangular(n, 8)
    sel = (np.random.default_rng(3).random(n) < 0.3).astype(np.int32) if howmny == "S" else None
    check_cplx_parity(S, P, al, be, side, howmny, sel, cond=True)


def test_complex64_interface():
    n = 60
    S, P, al, be = G.reference_style_complex(n, 77)
    gL, gR, gm, gi = run_cplx("gpu", "B", "A", None, S, P, al, be, prefix="c")
    oL, oR, om, oi = run_cplx("oracle", "B", "A", None, S, P, al, be, prefix="z")
    assert (gm, gi) == (om, oi)
    O.normalize(oR)
    O.normalize(oL)
    assert np.max(np.abs(gR - oR)) < 1e-4 and np.max(np.abs(gL - oL)) < 1e-4


def test_float_interface():
    import xtgevc3_amd as X

    n = 70
    S, P, ar, ai, be = G.reference_style_real(n, 78)
    VL = np.zeros((n, n), dtype=np.float32, order="F")
    VR = np.zeros((n, n), dtype=np.float32, order="F")
    f = np.float32
    m, info, _ = X.stgevc3("B", "A", None, np.asfortranarray(S.astype(f)), np.asfortranarray(P.astype(f)),
                           ar.astype(f), ai.astype(f), be.astype(f), VL=VL, VR=VR)
    oL = np.zeros((n, n), order="F")
    oR = np.zeros((n, n), order="F")
    S32, P32 = S.astype(f).astype(np.float64), P.astype(f).astype(np.float64)
    O.tgevc3("d", "B", "A", None, S32, P32, ar.astype(f).astype(float), be.astype(f).astype(float),
             alphai=ai.astype(f).astype(float), VL=oL, VR=oR)
    kind, idx = O.real_layout("A", None, ai.astype(f).astype(float))
    O.normalize(oL, kind)
    O.normalize(oR, kind)
    assert (m, info) == (n, 0)
    assert np.max(np.abs(VR - oR)) < 1e-4 and np.max(np.abs(VL - oL)) < 1e-4
