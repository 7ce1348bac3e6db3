"""Exact-arithmetic answers for the reference's 4x4 known-answer pencils
(tests/test_dtgevc3.cpp:322-326, :355-359, :388-395; tests/test_ztgevc3.cpp:158-184).

The eigenvectors are computed in exact (Gaussian) rationals from the defining equations
alone -- x in the null space of beta*S - alpha*P with x_k = 1 below/above the diagonal
block, and for a complex pair the phase fixed by the null vector of the 2x2 block's first
row (right) or first column (left) -- then normalised and compared with the oracle and,
when present, the compiled reference library.  This pins the oracle independently of any
floating-point implementation."""
import numpy as np
import pytest
import sympy as sp

from oracle import oracle as O

EPS = np.finfo(np.float64).eps

S_COLS = [1, 0, 0, 0, 2, 2, -5, 0, 3, 5, 2, 0, 4, -1, 2, 3]
P_COLS = [2, 0, 0, 0, 1, 1, 0, 0, -1, 0, 1, 0, 3, 2, -1, 2]


def _mat(cols):
    return sp.Matrix(4, 4, lambda r, c: sp.Integer(cols[r + 4 * c]))


def _exact_right(S, P, al, be, k, w, conv_first_row=True):
    """Right eigenvector (column vector) for eigenvalue al/be of the block starting at k."""
    n = S.shape[0]
    C = be * S - al * P
    x = sp.zeros(n, 1)
    if w == 1:
        x[k] = 1
    else:
        # null vector of the first row of the 2x2 block: (-c12, c11)
        x[k] = -C[k, k + 1]
        x[k + 1] = C[k, k]
    # back substitution on rows above the block (exact, any pivoting gives the same x)
    if k > 0:
        T = C[:k, :k]
        rhs = -(C[:k, k:k + w] * x[k:k + w, 0])
        x[:k, 0] = T.LUsolve(rhs)
    return x


def _exact_left(S, P, al, be, k, w):
    """Left eigenvector y with y^H (be S - al P) = 0: phase from the first column of the
    2x2 block, y_k = -a s21, y_{k+1} = a s11 - conj(b) p11 (a = be, b = al)."""
    n = S.shape[0]
    C = be * S - al * P
    CH = C.H
    y = sp.zeros(n, 1)
    if w == 1:
        y[k] = 1
    else:
        y[k] = -be * S[k + 1, k]
        y[k + 1] = be * S[k, k] - sp.conjugate(al) * P[k, k]
    if k + w < n:
        T = CH[k + w:, k + w:]
        rhs = -(CH[k + w:, k:k + w] * y[k:k + w, 0])
        y[k + w:, 0] = T.LUsolve(rhs)
    return y


def _to_columns(vecs, kinds, n):
    """complex exact vectors -> real-interface columns (pairs: re, im)."""
    out = np.zeros((n, sum(2 if kd else 1 for kd in kinds)))
    c = 0
    for v, kd in zip(vecs, kinds):
        z = np.array([complex(sp.N(sp.re(e), 40)) + 1j * complex(sp.N(sp.im(e), 40)) for e in v], dtype=complex)
        if kd == 0:
            out[:, c] = z.real
            c += 1
        else:
            out[:, c] = z.real
            out[:, c + 1] = z.imag
            c += 2
    return out


def _kat(scale=1, p33=2, b3=2):
    S = _mat(S_COLS) * scale
    Pc = list(P_COLS)
    Pc[15] = p33
    P = _mat(Pc) * scale
    al = [sp.Integer(1) * scale, (2 + 5 * sp.I) * scale, None, sp.Integer(3) * scale]
    be = [sp.Integer(2) * scale, sp.Integer(1) * scale, None, sp.Integer(b3) * scale]
    return S, P, al, be


CASES = {"plain": dict(), "infinite": dict(p33=0, b3=0)}


@pytest.mark.parametrize("case", ["plain", "infinite", "big", "small"])
def test_real_kat_exact(case):
    if case in CASES:
        S, P, al, be = _kat(**CASES[case])
        s = 1.0
    else:
        S, P, al, be = _kat()
        s = 1e150 if case == "big" else 1e-150
    units = [(0, 1), (1, 2), (3, 1)]
    right = [_exact_right(S, P, al[k], be[k], k, w) for k, w in units]
    left = [_exact_left(S, P, al[k], be[k], k, w) for k, w in units]
    kinds = [0, 1, 0]
    kind = np.array([0, 1, 2, 0], dtype=np.int32)
    eR = O.normalize(np.asfortranarray(_to_columns(right, kinds, 4)), kind)
    eL = O.normalize(np.asfortranarray(_to_columns(left, kinds, 4)), kind)

    Sf = np.asfortranarray(np.array(S, dtype=np.float64) * s)
    Pf = np.asfortranarray(np.array(P, dtype=np.float64) * s)
    ar = np.array([1.0, 2.0, 2.0, 3.0]) * s
    ai = np.array([0.0, 5.0, -5.0, 0.0]) * s
    bb = np.array([2.0, 1.0, 1.0, 0.0 if case == "infinite" else 2.0]) * s
    impls = ["oracle"] + (["ref"] if O.have_ref() else [])
    for impl in impls:
        VL = np.asfortranarray(np.zeros((4, 4)))
        VR = np.asfortranarray(np.zeros((4, 4)))
        m, info, _ = O.tgevc3("d", "B", "A", None, Sf, Pf, ar, bb, alphai=ai, VL=VL, VR=VR, impl=impl)
        assert (m, info) == (4, 0)
        O.normalize(VL, kind)
        O.normalize(VR, kind)
        assert np.max(np.abs(VR - eR)) < 100 * EPS, (impl, VR, eR)
        assert np.max(np.abs(VL - eL)) < 100 * EPS, (impl, VL, eL)


def test_complex_kat_exact():
    I = sp.I
    S = sp.zeros(4, 4)
    P = sp.zeros(4, 4)
    ent = {(0, 0): (1 + I, 2), (0, 1): (2, 1 + I), (1, 1): (2 - I, 1 - I), (0, 2): (-5 + I, 0), (1, 2): (2 * I, -1),
           (2, 2): (3, 1 + I), (0, 3): (5, 0), (1, 3): (2 + I, 2 * I), (2, 3): (4 - I, 3), (3, 3): (3 + 2 * I, 2 - I)}
    for (r, c), (s_, p_) in ent.items():
        S[r, c] = s_
        P[r, c] = p_
    right, left = [], []
    for k in range(4):
        al, be = S[k, k], P[k, k]
        C = be * S - al * P
        x = sp.zeros(4, 1)
        x[k] = 1
        if k > 0:
            x[:k, 0] = C[:k, :k].LUsolve(-C[:k, k])
        right.append(x)
        y = sp.zeros(4, 1)
        y[k] = 1
        CH = C.H
        if k < 3:
            y[k + 1:, 0] = CH[k + 1:, k + 1:].LUsolve(-CH[k + 1:, k])
        left.append(y)

    def cols(vs):
        return np.asfortranarray(np.array([[complex(sp.N(e, 40)) for e in v] for v in vs]).T)

    eR = O.normalize(cols(right))
    eL = O.normalize(cols(left))
    Sf = np.asfortranarray(np.array(S.evalf(), dtype=np.complex128))
    Pf = np.asfortranarray(np.array(P.evalf(), dtype=np.complex128))
    al = np.diag(Sf).copy()
    be = np.diag(Pf).copy()
    for impl in ["oracle"] + (["ref"] if O.have_ref() else []):
        VL = np.zeros((4, 4), dtype=np.complex128, order="F")
        VR = np.zeros((4, 4), dtype=np.complex128, order="F")
        m, info, _ = O.tgevc3("z", "B", "A", None, Sf, Pf, al, be, VL=VL, VR=VR, impl=impl)
        assert (m, info) == (4, 0)
        O.normalize(VL)
        O.normalize(VR)
        assert np.max(np.abs(VR - eR)) < 100 * EPS
        assert np.max(np.abs(VL - eL)) < 100 * EPS
