"""CPU tests of the boundary: libxtgevc3.so loads and exports every entry point the headers
declare, and the host-side argument semantics (m, info, lwork query) equal the reference's,
checked by brute force against the oracle (which restates src/dtgevc3.cpp:1173-1242 and
src/ztgevc3.cpp:456-516) on paths that return before any device work."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import xtgevc3_amd as X
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_c_functions():
    text = open(os.path.join(ROOT, "include", "xtgevc3.h")).read()
    return sorted(set(re.findall(r"\b(xtgevc3_\w+)\s*\(", text)))


def test_library_exports_every_c_symbol():
    names = _declared_c_functions()
    assert names, "no declarations found"
    lib = X.lib()
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(X.EXPORTED_SYMBOLS)
    assert lib.xtgevc3_abi_version() == 1


def test_library_exports_cxx_dropins():
    out = subprocess.run(["nm", "-DC", "--defined-only", X.LIB_PATH], capture_output=True, text=True, check=True).stdout
    text = open(os.path.join(ROOT, "include", "tgevc3.h")).read()
    for fn in re.findall(r"^void (\w+)\(", text, flags=re.M):
        assert re.search(r"\b" + fn + r"\(char, char, int const\*, int,", out), fn
    # the drop-in header keeps the reference's parameter lists (include/tgevc3.h:10-13)
    assert "dtgevc3(char, char, int const*, int, double const*, int, double const*, int, double const*, " \
           "double const*, double const*, double*, int, double*, int, int, int*, double*, int, int*)" in out


def _both_real(prefix, side, howmny, sel, n, ai, mm, lwork):
    dt = np.float64 if prefix == "d" else np.float32
    S = np.asfortranarray(np.triu(np.ones((n, n))).astype(dt))
    P = S.copy(order="F")
    ar = np.ones(n, dtype=dt)
    be = np.ones(n, dtype=dt)
    res = []
    for impl in ("gpu", "oracle"):
        VL = np.zeros((n, max(mm, 1)), dtype=dt, order="F")
        VR = np.zeros((n, max(mm, 1)), dtype=dt, order="F")
        work = np.full(max(lwork, 1), -7, dtype=dt)
        m = ctypes.c_int(-5)
        info = ctypes.c_int(-99)
        if impl == "gpu":
            fn = getattr(X.lib(), f"xtgevc3_{prefix}tgevc3")
        else:
            fn = getattr(O._load(O.ORACLE_SO), f"oracle_{prefix}tgevc3")
        c = ctypes
        fn(c.c_char(side.encode()), c.c_char(howmny.encode()), None if sel is None else O._ptr(sel), c.c_int(n),
           O._ptr(S), c.c_int(max(n, 1)), O._ptr(P), c.c_int(max(n, 1)), O._ptr(ar), O._ptr(ai.astype(dt)),
           O._ptr(be), O._ptr(VL), c.c_int(max(n, 1)), O._ptr(VR), c.c_int(max(n, 1)), c.c_int(mm), c.byref(m),
           O._ptr(work), c.c_int(lwork), c.byref(info))
        res.append((m.value, info.value, float(work[0])))
    return res


def _req(n, b=32):
    return 2 * n * (b + 1) + 4 * (b + 1) ** 2 + 2 * (b + 1)


@pytest.mark.parametrize("prefix", ["d", "s"])
def test_real_semantics_bruteforce(prefix):
    rng = np.random.default_rng(0)
    for trial in range(300):
        n = int(rng.integers(0, 9))
        ai = np.zeros(max(n, 1))
        k = 0
        while k < n:
            if k + 1 < n and rng.random() < 0.4:
                ai[k], ai[k + 1] = 1.0, -1.0
                k += 2
            else:
                if rng.random() < 0.1:
                    ai[k] = 2.0  # lone nonzero alphai at the last index: treated as real
                k += 1
        howmny = str(rng.choice(["A", "B", "S", "X"]))
        sel = (rng.random(max(n, 1)) < 0.5).astype(np.int32) if rng.random() < 0.8 else None
        side = str(rng.choice(["R", "L", "B", "Q"]))
        case = rng.integers(0, 4)
        # force an early return so that no device work happens: query, short lwork,
        # mm < m, bad howmny or n == 0
        if case == 0:
            mm, lwork = n, -1
        elif case == 1:
            mm, lwork = n, int(rng.integers(0, max(1, 4 * n + 20)))
        elif case == 2:
            mm, lwork = -1, _req(n)
        else:
            mm, lwork = n, _req(n)
            howmny = "X"
        if n == 0 and case == 0:
            pass
        g, o = _both_real(prefix, side, howmny, sel, n, ai[:max(n, 1)], mm, lwork)
        assert g == o, (trial, n, howmny, side, case, g, o)


def test_select_count_matches_reference_rule():
    """m for random selections and pair layouts (pairs count twice; a pair is selected when
    either of its columns is; alphai[n-1] != 0 alone is a real eigenvalue)."""
    rng = np.random.default_rng(1)
    for trial in range(200):
        n = int(rng.integers(1, 30))
        ai = np.zeros(n)
        k = 0
        while k < n:
            if k + 1 < n and rng.random() < 0.5:
                ai[k], ai[k + 1] = 1.0, -1.0
                k += 2
            else:
                k += 1
        if rng.random() < 0.2:
            ai[n - 1] = 3.0
        sel = (rng.random(n) < 0.5).astype(np.int32)
        g, o = _both_real("d", "R", "S", sel, n, ai, -1, _req(n))
        assert g == o and g[1] == -16
        # brute force from the rule itself
        m = 0
        k = 0
        while k < n:
            pair = ai[k] != 0 and k + 1 < n
            if pair:
                m += 2 if (sel[k] or sel[k + 1]) else 0
                k += 2
            else:
                m += 1 if sel[k] else 0
                k += 1
        assert g[0] == m


def test_complex_semantics_bruteforce():
    rng = np.random.default_rng(2)
    for trial in range(200):
        n = int(rng.integers(0, 9))
        howmny = str(rng.choice(["A", "B", "S", "X"]))
        sel = (rng.random(max(n, 1)) < 0.5).astype(np.int32) if rng.random() < 0.8 else None
        case = rng.integers(0, 3)
        mm, lwork = (n, -1) if case == 0 else ((n, int(rng.integers(0, 4 * n + 20))) if case == 1 else (-1, _req(n)))
        S = np.asfortranarray(np.eye(max(n, 1), dtype=np.complex128))[:n, :n].copy(order="F")
        res = []
        for impl in ("gpu", "oracle"):
            work = np.full(max(lwork, 1), -7, dtype=np.complex128)
            m = ctypes.c_int(-5)
            info = ctypes.c_int(-99)
            fn = X.lib().xtgevc3_ztgevc3 if impl == "gpu" else O._load(O.ORACLE_SO).oracle_ztgevc3
            al = np.ones(max(n, 1), dtype=np.complex128)
            V = np.zeros((max(n, 1), max(n, 1)), dtype=np.complex128, order="F")
            c = ctypes
            fn(c.c_char(b"B"), c.c_char(howmny.encode()), None if sel is None else O._ptr(sel), c.c_int(n),
               O._ptr(S), c.c_int(max(n, 1)), O._ptr(S), c.c_int(max(n, 1)), O._ptr(al), O._ptr(al), O._ptr(V),
               c.c_int(max(n, 1)), O._ptr(V), c.c_int(max(n, 1)), c.c_int(mm), c.byref(m), O._ptr(work),
               c.c_int(lwork), c.byref(info))
            res.append((m.value, info.value, complex(work[0])))
        assert res[0] == res[1], (trial, n, howmny, case, res)


def test_python_mirror_query_idiom():
    n = 12
    S = np.asfortranarray(np.triu(np.ones((n, n))))
    ai = np.zeros(n)
    m, info, w0 = X.dtgevc3("R", "A", None, S, S, np.ones(n), ai, np.ones(n), VR=np.zeros((n, n), order="F"),
                            lwork=-1)
    assert (info, int(w0)) == (0, _req(n))
    m, info, _ = X.dtgevc3("R", "Z", None, S, S, np.ones(n), ai, np.ones(n), VR=np.zeros((n, n), order="F"))
    assert info == -2


@pytest.mark.parametrize("precision", [2, 4, -1])
def test_unknown_precision_is_rejected_before_device_work(precision):
    """Precision selectors other than FP64, FP32 products and split-int8 return
    XTGEVC3_INFO_UNSUPPORTED (also through info) without touching any pointer or the GPU."""
    c = ctypes
    lib = X.lib()
    info = c.c_int(-99)
    m = c.c_int(-5)
    rc = lib.xtgevc3_dtgevc3_device_ex(c.c_char(b"R"), c.c_char(b"A"), None, c.c_int(4), None, c.c_int(4), None,
                                       c.c_int(4), None, None, None, None, c.c_int(4), None, c.c_int(4), c.c_int(4),
                                       c.byref(m), c.byref(info), None, c.c_int(precision))
    assert rc == X.INFO_UNSUPPORTED and info.value == X.INFO_UNSUPPORTED
    assert "precision" in X.last_error()
