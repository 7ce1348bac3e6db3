"""The drop-in boundary called as the reference's C++ callers call it: tests/cxx/dropin_caller
includes include/tgevc3.h, links libxtgevc3.so and calls the C++-linkage (mangled) dtgevc3 and
ztgevc3 with the two-call lwork idiom of tests/test_lapack_d.cpp:278-286.

* CPU: the argument checks and early returns (lwork query, bad howmny, mm < m, short lwork,
  n = 0, select = NULL with howmny = 'S') print the same m / info / work[0] as the same program
  linked against the reference library compiled from its sources (oracle/_ref), or, where that
  is not built, the reference's documented values.
* GPU: side = howmny = 'B' at n = 100 (real and complex), checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import EPS, compare_sampled, eigvec_cond_sampled

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CALLER = os.path.join(HERE, "cxx", "dropin_caller")
REF_CALLER = os.path.join(ROOT, "oracle", "_ref", "dropin_caller_ref")

# lwork formula 2n(b+1) + 4(b+1)^2 + 2(b+1) with b = 32 at n = 8 (src/dtgevc3.cpp:1219-1242)
DOCUMENTED = """d_query m=8 info=0 work0=4950
d_bad_howmny m=-7 info=-2 work0=0
d_mm_short m=8 info=-16 work0=0
d_lwork_short m=8 info=-19 work0=0
d_n0 m=0 info=0 work0=0
d_select_null m=0 info=0 work0=0
z_query m=8 info=0 work0=4950
z_bad_howmny m=-7 info=-2 work0=0
z_mm_short m=8 info=-16 work0=0
z_lwork_short m=8 info=-19 work0=0
"""


def _caller():
    if not os.path.exists(CALLER):
        pytest.fail(f"{CALLER} is not built (make -C tests/cxx, part of __graft_entry__.build())")
    return CALLER


def test_dropin_cxx_early_returns():
    out = subprocess.run([_caller(), "early"], check=True, capture_output=True, text=True).stdout
    if os.path.exists(REF_CALLER):
        ref = subprocess.run([REF_CALLER, "early"], check=True, capture_output=True, text=True).stdout
        assert out == ref
    assert out == DOCUMENTED


def _read(path, prefix):
    raw = open(path, "rb").read()
    n, m, info = np.frombuffer(raw[:12], dtype=np.int32)
    off = 12
    dt = np.float64 if prefix == "d" else np.complex128

    def take(count, shape=None):
        nonlocal off
        a = np.frombuffer(raw, dtype=dt, count=count, offset=off).copy()
        off += a.nbytes
        return a if shape is None else a.reshape(shape, order="F")

    S, P = take(n * n, (n, n)), take(n * n, (n, n))
    ev = [take(n) for _ in range(3 if prefix == "d" else 2)]
    VL, VR = take(n * n, (n, n)), take(n * n, (n, n))
    return int(n), int(m), int(info), S, P, ev, VL, VR


@pytest.mark.gpu
@pytest.mark.parametrize("prefix", ["d", "z"])
def test_dropin_cxx_solve_matches_oracle(prefix, tmp_path):
    path = str(tmp_path / f"out_{prefix}.bin")
    subprocess.run([_caller(), "solve", prefix, "100", path], check=True, timeout=120)
    n, m, info, S, P, ev, VL, VR = _read(path, prefix)
    eye = np.eye(n, dtype=S.dtype)
    oL, oR = np.asfortranarray(eye.copy()), np.asfortranarray(eye.copy())
    if prefix == "d":
        ar, ai, be = ev
        om, oinfo, _ = O.tgevc3("d", "B", "B", None, S, P, ar, be, alphai=ai, VL=oL, VR=oR)
        kind, idx = O.real_layout("B", None, ai)
        al, bb = O.eig_alpha_beta(ar, ai, be, idx)
        O.normalize(oL, kind)
        O.normalize(oR, kind)
    else:
        al, bb = ev
        om, oinfo, _ = O.tgevc3("z", "B", "B", None, S, P, al, bb, VL=oL, VR=oR)
        kind, idx = np.zeros(n, dtype=np.int32), np.arange(n)
        O.normalize(oL)
        O.normalize(oR)
    assert (m, info) == (om, oinfo) == (n, 0)
    for g, o, left in ((VL, oL, True), (VR, oR, False)):
        assert np.all(np.isfinite(g))
        cond = eigvec_cond_sampled(S, P, al, bb, kind, idx, left=left)
        bad, err, tol = compare_sampled(g, o, cond)
        assert bad.size == 0, f"cols {bad[:8]} err {err[bad[:8]]} tol {tol[bad[:8]]}"
