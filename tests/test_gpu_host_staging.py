"""Host-pointer entry points on matrices with padded leading dimensions, several staging chunks
long.  The vectors must equal the device-pointer path's bit for bit (rounded to FP32 / complex64
for stgevc3 / ctgevc3, whose matrices are widened on the device), and rows past n, and columns
past m, of the caller's arrays must be left untouched."""
import ctypes

import numpy as np
import pytest

import xtgevc3_amd as X
from xtgevc3_amd import pencils as G

pytestmark = pytest.mark.gpu

N = 3000  # 3 to 6 chunks of 32 MiB per matrix, depending on the element type
PAD = 5
SENT = 7.25
WIDE = {"d": np.float64, "s": np.float64, "z": np.complex128, "c": np.complex128}
HOST = {"d": np.float64, "s": np.float32, "z": np.complex128, "c": np.complex64}


def _padded(a, dt):
    out = np.full((a.shape[0] + PAD, a.shape[1]), SENT, dtype=dt, order="F")
    out[: a.shape[0]] = a
    return out


def _host_call(prefix, side, howmny, sel, S, P, ev, VL, VR, mm):
    """The reference's two-call idiom through the C ABI with the padded arrays' leading dimension."""
    real = prefix in ("d", "s")
    dt = HOST[prefix]
    fn = getattr(X.lib(), f"xtgevc3_{prefix}tgevc3")
    alpha, alphai, beta = (ev[0], ev[1], ev[2]) if real else (ev[0], None, ev[1])
    m, info = ctypes.c_int(-1), ctypes.c_int(-99)
    q = np.zeros(1, dtype=dt)
    X._invoke(fn, real, side, howmny, sel, N, S, P, alpha, alphai, beta, dt, VL, VR, mm, m, q, -1, info)
    assert info.value == 0
    work = np.zeros(max(int(np.real(q[0])), 1), dtype=dt)
    X._invoke(fn, real, side, howmny, sel, N, S, P, alpha, alphai, beta, dt, VL, VR, mm, m, work, len(work), info)
    assert info.value == 0, X.last_error()
    return m.value


@pytest.mark.parametrize("side,howmny", [("B", "B"), ("R", "S")])
@pytest.mark.parametrize("prefix", ["d", "s", "z", "c"])
def test_host_staging_padded(prefix, side, howmny):
    real = prefix in ("d", "s")
    dt = HOST[prefix]
    rng = np.random.default_rng(11)
    if real:
        S, P, ar, ai, be = G.real_quasi_triangular(N, 21)
        ev = [ar.astype(dt), ai.astype(dt), be.astype(dt)]
    else:
        S, P, al, be = G.complex_triangular(N, 22)
        ev = [al.astype(dt), be.astype(dt)]
    S, P = S.astype(dt), P.astype(dt)
    sel = None
    if howmny == "S":
        if real:
            # every selected pair lists both of its columns
            sel = np.zeros(N, dtype=np.int32)
            k = 0
            while k < N:
                w = 2 if ev[1][k] != 0 else 1
                if rng.random() < 0.3:
                    sel[k : k + w] = 1
                k += w
        else:
            sel = (rng.random(N) < 0.3).astype(np.int32)
    mm = N
    back = howmny == "B"

    def start():
        a = rng.standard_normal((N, mm))
        if not real:
            a = a + 1j * rng.standard_normal((N, mm))
        return a.astype(dt) if back else np.zeros((N, mm), dtype=dt)

    left = side in ("L", "B")
    VL0 = start() if left else None
    VR0 = start()

    # device-pointer path on the widened inputs
    wide = WIDE[prefix]
    dS, dP = X.fortran_device(S.astype(wide)), X.fortran_device(P.astype(wide))
    dVL = X.fortran_device(VL0.astype(wide)) if left else None
    dVR = X.fortran_device(VR0.astype(wide))
    wev = [e.astype(wide) for e in ev]
    if real:
        m_dev, info = X.tgevc3_device("d", side, howmny, sel, dS, dP, wev[0], wev[2], alphai=wev[1], VL=dVL, VR=dVR,
                                      mm=mm)
    else:
        m_dev, info = X.tgevc3_device("z", side, howmny, sel, dS, dP, wev[0], wev[1], VL=dVL, VR=dVR, mm=mm)
    assert info == 0
    expR = X.from_fortran_device(dVR)[:, :m_dev].astype(dt)
    expL = X.from_fortran_device(dVL)[:, :m_dev].astype(dt) if left else None

    # host-pointer path with padded leading dimensions
    hS, hP = _padded(S, dt), _padded(P, dt)
    hVL = _padded(VL0, dt) if left else None
    hVR = _padded(VR0, dt)
    m = _host_call(prefix, side, howmny, sel, hS, hP, ev, hVL, hVR, mm)
    assert m == m_dev
    for h, exp, init in ((hVR, expR, VR0), (hVL, expL, VL0)):
        if h is None:
            continue
        assert np.array_equal(h[:N, :m], exp)
        assert np.all(h[N:, :] == SENT)
        assert np.array_equal(h[:N, m:], init[:, m:])
    assert np.all(hS[N:] == SENT) and np.array_equal(hS[:N], S)


@pytest.mark.parametrize("side", ["R", "L"])
@pytest.mark.parametrize("prefix", ["d", "s", "z", "c"])
def test_host_back_transform_pieces_one_side(prefix, side):
    """The host entries copy their back-transformed vectors out in pieces per side (right
    pieces first, then left; rounded to FP32 on the device for s / c); one side alone takes
    only that side's pieces and events."""
    test_host_staging_padded(prefix, side, "B")
