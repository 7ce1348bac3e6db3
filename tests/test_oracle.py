"""Pins the oracle (CPU restatement, oracle/tgevc3_oracle.cpp) before it is trusted:
  * against the golden fixtures written by the reference library itself (tests/golden),
  * against the reference library compiled in this container (oracle/_ref), when built,
  * against LAPACK xTGEVC (independent algorithm, same eigenvectors up to scale / phase),
  * against residual bounds (the pencil equation itself).
A planted mistake in the oracle (dropped term, wrong sign, transposed operand) fails at
least the golden and LAPACK comparisons."""
import ctypes
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O
from xtgevc3_amd import pencils as G

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EPS = np.finfo(np.float64).eps


def _golden(name):
    return dict(np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False))


def _golden_names(prefix):
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLD, prefix + "*.npz")))


def _rel(a, b):
    scale = max(np.max(np.abs(b)), np.finfo(np.float64).tiny)
    return np.max(np.abs(a - b)) / scale


@pytest.mark.parametrize("name", _golden_names("kat_d") + _golden_names("rand_d") + _golden_names("scale_d"))
def test_oracle_real_matches_golden_raw(name):
    g = _golden(name)
    n = g["S"].shape[0]
    howmny = str(g["howmny"])
    sel = g["sel"].astype(np.int32) if "sel" in g and g["sel"].size else None
    vin = g.get("Vin", np.eye(n))
    VL = np.asfortranarray(vin.copy())
    VR = np.asfortranarray(vin.copy())
    m, info, _ = O.tgevc3("d", str(g["side"]), howmny, sel, g["S"], g["P"], g["ar"], g["be"], alphai=g["ai"],
                          VL=VL, VR=VR)
    assert (m, info) == (int(g["m"]), int(g["info"]))
    # raw vectors, scale included: same algorithm, BLAS rounding aside
    for mine, ref in ((VL, g["VL"]), (VR, g["VR"])):
        for c in range(m):
            assert _rel(mine[:, c], ref[:, c]) < 1e-9, (name, c)


@pytest.mark.parametrize("name", _golden_names("kat_z") + _golden_names("rand_z") + _golden_names("scale_z"))
def test_oracle_complex_matches_golden_raw(name):
    g = _golden(name)
    n = g["S"].shape[0]
    sel = g["sel"].astype(np.int32) if "sel" in g and g["sel"].size else None
    vin = g.get("Vin", np.eye(n, dtype=np.complex128))
    VL = np.asfortranarray(vin.astype(np.complex128).copy())
    VR = np.asfortranarray(vin.astype(np.complex128).copy())
    m, info, _ = O.tgevc3("z", str(g["side"]), str(g["howmny"]), sel, g["S"], g["P"], g["al"], g["be"], VL=VL, VR=VR)
    assert (m, info) == (int(g["m"]), int(g["info"]))
    for mine, ref in ((VL, g["VL"]), (VR, g["VR"])):
        for c in range(m):
            assert _rel(mine[:, c], ref[:, c]) < 1e-9, (name, c)


@pytest.mark.parametrize("name", ["kat_s_big", "kat_s_small", "kat_s_inf", "kat_s_sub"])
def test_oracle_float_matches_golden(name):
    g = _golden(name)
    VL = np.asfortranarray(np.eye(4, dtype=np.float32))
    VR = np.asfortranarray(np.eye(4, dtype=np.float32))
    m, info, _ = O.tgevc3("s", "B", "B", None, g["S"], g["P"], g["ar"], g["be"], alphai=g["ai"], VL=VL, VR=VR)
    assert (m, info) == (int(g["m"]), int(g["info"]))
    assert _rel(VL, g["VL"]) < 1e-5 and _rel(VR, g["VR"]) < 1e-5


def test_oracle_complex64_subnormal_beta_matches_golden():
    """complex64 KAT with beta_3 = 1e-40: the reference (ctgevc3) counts it as infinite."""
    g = _golden("kat_c_sub")
    VL = np.asfortranarray(np.eye(4, dtype=np.complex64))
    VR = np.asfortranarray(np.eye(4, dtype=np.complex64))
    m, info, _ = O.tgevc3("c", "B", "B", None, g["S"], g["P"], g["al"], g["be"], VL=VL, VR=VR)
    assert (m, info) == (int(g["m"]), int(g["info"]))
    assert _rel(VL, g["VL"]) < 1e-5 and _rel(VR, g["VR"]) < 1e-5


def _ref_or_skip():
    if not O.have_ref():
        pytest.skip("reference library not built here (needs /root/reference)")


@pytest.mark.parametrize("prefix", ["d", "s", "z", "c"])
@pytest.mark.parametrize("n", [1, 3, 31, 32, 33, 34, 65, 100])
def test_oracle_matches_reference_library(prefix, n):
    _ref_or_skip()
    rng = np.random.default_rng(1000 + n)
    for side in ("R", "L", "B"):
        for howmny in ("A", "B", "S"):
            sel = (rng.random(n) < 0.5).astype(np.int32) if howmny == "S" else None
            if prefix in "ds":
                S, P, ar, ai, be = G.reference_style_real(n, 7 * n + 1)
                dt = np.float64 if prefix == "d" else np.float32
                vin = G.random_orthogonal(n, 5) if howmny == "B" else np.zeros((n, n))
                outs = []
                for impl in ("oracle", "ref"):
                    VL = np.asfortranarray(vin.astype(dt))
                    VR = np.asfortranarray(vin.astype(dt))
                    r = O.tgevc3(prefix, side, howmny, sel, S.astype(dt), P.astype(dt), ar.astype(dt), be.astype(dt),
                                 alphai=ai.astype(dt), VL=VL, VR=VR, impl=impl)
                    outs.append((VL, VR, r[0], r[1]))
            else:
                S, P, al, be = G.reference_style_complex(n, 7 * n + 2)
                dt = np.complex128 if prefix == "z" else np.complex64
                vin = G.random_unitary(n, 6) if howmny == "B" else np.zeros((n, n))
                outs = []
                for impl in ("oracle", "ref"):
                    VL = np.asfortranarray(vin.astype(dt))
                    VR = np.asfortranarray(vin.astype(dt))
                    r = O.tgevc3(prefix, side, howmny, sel, S.astype(dt), P.astype(dt), al.astype(dt), be.astype(dt),
                                 VL=VL, VR=VR, impl=impl)
                    outs.append((VL, VR, r[0], r[1]))
            (aL, aR, am, ai_), (bL, bR, bm, bi) = outs
            assert (am, ai_) == (bm, bi)
            tol = 1e-9 if prefix in "dz" else 2e-3
            for x, y in ((aL, bL), (aR, bR)):
                for c in range(am):
                    assert _rel(x[:, c], y[:, c]) < tol, (prefix, n, side, howmny, c)


def _lapack_dtgevc(S, P):
    lib = O._openblas()
    n = S.shape[0]
    VL = np.zeros((n, n), order="F")
    VR = np.zeros((n, n), order="F")
    work = np.zeros(6 * n + 8)
    m = ctypes.c_int(0)
    info = ctypes.c_int(0)
    ci = lambda v: ctypes.byref(ctypes.c_int(v))
    Sf = np.asfortranarray(S)
    Pf = np.asfortranarray(P)
    lib.scipy_dtgevc_(ctypes.c_char_p(b"B"), ctypes.c_char_p(b"A"), None, ci(n), O._ptr(Sf), ci(n), O._ptr(Pf),
                      ci(n), O._ptr(VL), ci(n), O._ptr(VR), ci(n), ci(n), ctypes.byref(m), O._ptr(work),
                      ctypes.byref(info), ctypes.c_size_t(1), ctypes.c_size_t(1))
    assert info.value == 0
    return VL, VR


def _phase_invariant_err(x, y):
    """max|c x - y| / max|y| with c the least-squares complex factor (x^H y)/(x^H x): the
    two vectors span the same line whatever phase convention each one follows."""
    c = np.vdot(x, y) / np.vdot(x, x)
    return np.max(np.abs(c * x - y)) / np.max(np.abs(y))


@pytest.mark.parametrize("n", [5, 40, 90])
def test_oracle_matches_lapack_directions(n):
    S, P, Q, Z, ar, ai, be = G.qz_pencil(n, 300 + n)
    kind, idx = O.real_layout("A", None, ai)
    al, bb = O.eig_alpha_beta(ar, ai, be, idx)
    VL = np.zeros((n, n), order="F")
    VR = np.zeros((n, n), order="F")
    O.tgevc3("d", "B", "A", None, S, P, ar, be, alphai=ai, VL=VL, VR=VR)
    O.normalize(VL, kind)
    O.normalize(VR, kind)
    LL, LR = _lapack_dtgevc(S, P)
    for left, mine, lap in ((True, VL, LL), (False, VR, LR)):
        ck = O.eigvec_cond(S, P, al, bb, kind, idx, left=left)
        a = O.complex_columns(mine, kind)
        b = O.complex_columns(lap, kind)
        for c in range(n):
            err = _phase_invariant_err(a[:, c], b[:, c])
            assert err < 1000 * EPS * max(ck[c], 1.0), (left, c, err, ck[c])
            if kind[c] == 0:
                # real vectors: same sign convention up to a positive factor
                assert np.max(np.abs(mine[:, c] - lap[:, c])) < 1000 * EPS * max(ck[c], 1.0)


@pytest.mark.parametrize("n", [4, 50, 120])
def test_oracle_residuals(n):
    S, P, ar, ai, be = G.real_quasi_triangular(n, 400 + n)
    kind, idx = O.real_layout("A", None, ai)
    al, bb = O.eig_alpha_beta(ar, ai, be, idx)
    VL = np.zeros((n, n), order="F")
    VR = np.zeros((n, n), order="F")
    O.tgevc3("d", "B", "A", None, S, P, ar, be, alphai=ai, VL=VL, VR=VR)
    for left, V in ((False, VR), (True, VL)):
        r = O.residual_B(S, P, O.complex_columns(V, kind), al, bb, left=left)
        assert r.max() < 10 * n, r.max()
    Sz, Pz, alz, bz = G.complex_triangular(n, 500 + n)
    VL = np.zeros((n, n), dtype=np.complex128, order="F")
    VR = np.zeros((n, n), dtype=np.complex128, order="F")
    O.tgevc3("z", "B", "A", None, Sz, Pz, alz, bz, VL=VL, VR=VR)
    for left, V in ((False, VR), (True, VL)):
        r = O.residual_B(Sz, Pz, V, alz, bz, left=left)
        assert r.max() < 10 * n, r.max()


def test_oracle_normalisation_convention():
    V = np.asfortranarray(np.array([[2.0, 1.0, -3.0], [-4.0, 1.0, 0.5], [1.0, -2.0, 0.0]]))
    kind = np.array([0, 1, 2], dtype=np.int32)
    O.normalize(V, kind)
    assert np.max(np.abs(V[:, 0])) == 1.0 and V[1, 0] == -1.0
    assert np.isclose(np.max(np.abs(V[:, 1]) + np.abs(V[:, 2])), 1.0)


@pytest.mark.parametrize("seed", [1, 2])
def test_implicit_gradient_reference_matches_finite_differences(seed):
    """Pins oracle.implicit_grad_right (the CPU reference for the backward kernel) against
    central differences (h = 1e-6) of the oracle forward."""
    n = 20
    S, P, ar, ai, be = G.real_quasi_triangular(n, seed)
    real_k = [k for k in range(n) if ai[k] == 0 and (k == 0 or ai[k - 1] == 0)]
    sel = np.zeros(n, dtype=np.int32)
    sel[real_k[-4:]] = 1
    sel[real_k[1]] = 1

    def fwd(S_, P_):
        a, b = ar.copy(), be.copy()
        k = np.nonzero(sel)[0]
        a[k], b[k] = S_[k, k], P_[k, k]
        VR = np.zeros((n, n), order="F")
        O.tgevc3("d", "R", "S", sel, S_, P_, a, b, alphai=ai, VR=VR)
        kind, idx = O.real_layout("S", sel, ai)
        return O.normalize(VR[:, :len(idx)].copy(order="F"), kind), idx

    V, idx = fwd(S, P)
    W = np.random.default_rng(seed).standard_normal(V.shape)
    rS, rP = O.implicit_grad_right(S, P, idx, V, W)
    scale = max(np.abs(rS).max(), np.abs(rP).max())
    h = 1e-6
    rng = np.random.default_rng(seed + 7)
    entries = [(int(i), int(j)) for i, j in np.argwhere(np.triu(np.ones((n, n))) != 0)]
    entries = [entries[q] for q in rng.choice(len(entries), 40, replace=False)] + [(k, k) for k in idx]
    # below the band the forward reads nothing: the difference quotient and the reference are 0
    lower = [(int(i), int(j)) for i, j in np.argwhere(np.tril(np.ones((n, n)), -2) != 0)]
    entries += [lower[q] for q in rng.choice(len(lower), 10, replace=False)]
    for i, j in entries:
        for which in ("S", "P"):
            Mp = (S if which == "S" else P).copy()
            Mm = Mp.copy()
            Mp[i, j] += h
            Mm[i, j] -= h
            Lp = np.sum(fwd(Mp, P)[0] * W) if which == "S" else np.sum(fwd(S, Mp)[0] * W)
            Lm = np.sum(fwd(Mm, P)[0] * W) if which == "S" else np.sum(fwd(S, Mm)[0] * W)
            fd = (Lp - Lm) / (2 * h)
            ref = (rS if which == "S" else rP)[i, j]
            assert abs(fd - ref) <= 1e-7 * scale, (which, i, j, fd, ref)


def test_tgsna_e_diagonal_closed_form():
    """Diagonal pencil: e_k are the eigenvectors, s_k = hypot(S_kk, P_kk) whatever their scale."""
    rng = np.random.default_rng(5)
    n = 9
    d_s, d_p = rng.standard_normal(n), np.abs(rng.standard_normal(n)) + 0.1
    S, P = np.diag(d_s), np.diag(d_p)
    s = O.tgsna_e(S, P, 3.0 * np.eye(n), -0.5 * np.eye(n), np.zeros(n, dtype=np.int32))
    assert np.allclose(s, np.hypot(d_s, d_p), rtol=1e-15, atol=0)


@pytest.mark.parametrize("n", [6, 15])
def test_tgsna_e_matches_independent_eigensolver(n):
    """s_k from the oracle's real-layout vectors of (A, B) (QZ + tgevc3, back-transformed, pair
    formula) against s_k from LAPACK xGGEV's complex left/right eigenvectors (scipy.linalg.eig)
    with the complex formula |(u^H A v, u^H B v)| / (|u| |v|)."""
    import scipy.linalg as sl

    rng = np.random.default_rng(40 + n)
    A = rng.standard_normal((n, n))
    B = rng.standard_normal((n, n))
    S, P, al, be, Q, Z = sl.ordqz(A, B, sort="lhp", output="real")
    ar, ai, be = al.real.copy(), al.imag.copy(), be.real.copy()
    S, P, Q, Z = (np.asfortranarray(x) for x in (S, P, Q, Z))
    kind, idx = O.real_layout("A", None, ai)
    VL = np.asfortranarray(Q.copy())
    VR = np.asfortranarray(Z.copy())
    O.tgevc3("d", "B", "B", None, S, P, ar, be, alphai=ai, VL=VL, VR=VR)
    mine = O.tgsna_e(A, B, VL, VR, kind)
    # the same quantity in the Schur basis: the value does not depend on the basis
    VLs = np.zeros((n, n), order="F")
    VRs = np.zeros((n, n), order="F")
    O.tgevc3("d", "B", "A", None, S, P, ar, be, alphai=ai, VL=VLs, VR=VRs)
    assert np.allclose(O.tgsna_e(S, P, VLs, VRs, kind), mine, rtol=1e-10, atol=0)
    w, vl, vr = sl.eig(A, B, left=True, right=True)
    lam = (ar + 1j * ai)[idx] / be[idx]
    for c in range(n):
        k = int(np.argmin(np.abs(w - lam[c])))
        u, v = vl[:, k], vr[:, k]
        ref = np.hypot(abs(np.vdot(u, A @ v)), abs(np.vdot(u, B @ v))) / (np.linalg.norm(u) * np.linalg.norm(v))
        assert abs(mine[c] - ref) <= 1e-9 * ref, (c, mine[c], ref)


def test_tgsna_e_complex_matches_independent_eigensolver():
    """ztgsna (JOB='E') restatement on the oracle's back-transformed complex vectors against the
    same formula on LAPACK xGGEV's eigenvectors (scipy.linalg.eig), matched by eigenvalue."""
    import scipy.linalg as sl

    n = 12
    rng = np.random.default_rng(77)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    S, P, al, be, Q, Z = sl.ordqz(A, B, sort="lhp", output="complex")
    S, P, Q, Z = (np.asfortranarray(x) for x in (S, P, Q, Z))
    VL = np.asfortranarray(Q.copy())
    VR = np.asfortranarray(Z.copy())
    O.tgevc3("z", "B", "B", None, S, P, al, be, VL=VL, VR=VR)
    mine = O.tgsna_e_complex(A, B, VL, VR)
    w, vl, vr = sl.eig(A, B, left=True, right=True)
    lam = al / be
    for c in range(n):
        k = int(np.argmin(np.abs(w - lam[c])))
        u, v = vl[:, k], vr[:, k]
        ref = np.hypot(abs(np.vdot(u, A @ v)), abs(np.vdot(u, B @ v))) / (np.linalg.norm(u) * np.linalg.norm(v))
        assert abs(mine[c] - ref) <= 1e-9 * ref, (c, mine[c], ref)
