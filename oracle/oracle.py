"""TEST INFRASTRUCTURE ONLY -- Python side of the oracle.

Loads ``liboracle_tgevc3.so`` (our CPU restatement of the reference algorithm, built from
``tgevc3_oracle.cpp`` by ``oracle/Makefile``) and, when it has been built, the reference
library itself (``_ref/libref_tgevc3.so``), and wraps both behind the reference's own call
convention (``include/tgevc3.h:10-13``: side/howmny by value, column-major arrays, work/lwork
query, ``m`` and ``info`` out-parameters).

It also holds the residual metrics of the reference's tests:
  * metric A: per vector ||(beta S - alpha P) v||_2 / ((|beta| ||S||_F + |alpha| ||P||_F) ||v||_2)
    (tests/test_dtgevc3.cpp:29-223, tests/test_ztgevc3.cpp:29-124),
  * metric B: DGET52-style 1-norm residual in units of eps
    (tests/test_lapack_d.cpp:21-168, tests/test_lapack_z.cpp).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg import this
module, and only as the checker.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle_tgevc3.so")
REF_SO = os.path.join(HERE, "_ref", "libref_tgevc3.so")

_c_int_p = ctypes.POINTER(ctypes.c_int)
_libs = {}


def _load(path):
    if path not in _libs:
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run `make -C oracle oracle ref`)")
        _libs[path] = ctypes.CDLL(path)
    return _libs[path]


def have_ref():
    return os.path.exists(REF_SO)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


_REAL = {"d": np.float64, "s": np.float32}
_CPLX = {"z": np.complex128, "c": np.complex64}


def tgevc3(prefix, side, howmny, select, S, P, alpha, beta, alphai=None, VL=None, VR=None, mm=None,
           lwork=None, impl="oracle"):
    """Run ``{prefix}tgevc3`` of the oracle (impl='oracle') or of the compiled reference
    (impl='ref') with the reference's argument meaning.

    Real prefixes ('d','s') take alpha = alphar and alphai; complex ('z','c') take complex alpha.
    VL/VR are column-major (Fortran-ordered) arrays modified in place (they hold Q/Z on entry
    for howmny='B').  Returns (m, info, work0) where work0 is work[0] after the call
    (the workspace size for lwork=-1).
    """
    lib = _load(ORACLE_SO if impl == "oracle" else REF_SO)
    fn = getattr(lib, ("oracle_" if impl == "oracle" else "ref_") + prefix + "tgevc3")
    real = prefix in _REAL
    dt = _REAL[prefix] if real else _CPLX[prefix]
    n = S.shape[0]
    S = np.asfortranarray(S, dtype=dt)
    P = np.asfortranarray(P, dtype=dt)
    for V in (VL, VR):
        if V is not None:
            assert V.dtype == dt and V.flags.f_contiguous
    if mm is None:
        mm = n
    sel = None if select is None else np.ascontiguousarray(select, dtype=np.int32)
    m = ctypes.c_int(-1)
    info = ctypes.c_int(-99)
    if lwork is None:
        q = np.zeros(1, dtype=dt)
        _call(fn, real, side, howmny, sel, n, S, P, alpha, alphai, beta, dt, VL, VR, mm, m, q, -1, info)
        lwork = int(np.real(q[0]))
    if lwork == -1:
        work = np.zeros(1, dtype=dt)
    else:
        work = np.zeros(max(lwork, 1), dtype=dt)
    _call(fn, real, side, howmny, sel, n, S, P, alpha, alphai, beta, dt, VL, VR, mm, m, work, lwork, info)
    return m.value, info.value, work[0]


def _call(fn, real, side, howmny, sel, n, S, P, alpha, alphai, beta, dt, VL, VR, mm, m, work, lwork, info):
    lds = max(1, S.shape[0])
    ldvl = VL.shape[0] if VL is not None else 1
    ldvr = VR.shape[0] if VR is not None else 1
    if real:
        ar = np.ascontiguousarray(alpha, dtype=dt)
        ai = np.ascontiguousarray(alphai, dtype=dt)
        be = np.ascontiguousarray(beta, dtype=dt)
        fn(ctypes.c_char(side.encode()), ctypes.c_char(howmny.encode()), _ptr(sel), ctypes.c_int(n), _ptr(S),
           ctypes.c_int(lds), _ptr(P), ctypes.c_int(lds), _ptr(ar), _ptr(ai), _ptr(be), _ptr(VL),
           ctypes.c_int(ldvl), _ptr(VR), ctypes.c_int(ldvr), ctypes.c_int(mm), ctypes.byref(m), _ptr(work),
           ctypes.c_int(lwork), ctypes.byref(info))
    else:
        al = np.ascontiguousarray(alpha, dtype=dt)
        be = np.ascontiguousarray(beta, dtype=dt)
        fn(ctypes.c_char(side.encode()), ctypes.c_char(howmny.encode()), _ptr(sel), ctypes.c_int(n), _ptr(S),
           ctypes.c_int(lds), _ptr(P), ctypes.c_int(lds), _ptr(al), _ptr(be), _ptr(VL), ctypes.c_int(ldvl),
           _ptr(VR), ctypes.c_int(ldvr), ctypes.c_int(mm), ctypes.byref(m), _ptr(work), ctypes.c_int(lwork),
           ctypes.byref(info))


def real_layout(howmny, select, alphai):
    """Output column kinds (0 real, 1 pair real part, 2 pair imaginary part) and eigenvalue
    index per output column, by the selection rule of dtgevc3.cpp:1192-1207."""
    lib = _load(ORACLE_SO)
    dt = alphai.dtype
    n = alphai.shape[0]
    kind = np.zeros(max(n, 1), dtype=np.int32)
    idx = np.zeros(max(n, 1), dtype=np.int32)
    sel = None if select is None else np.ascontiguousarray(select, dtype=np.int32)
    fn = lib.oracle_dselect_layout if dt == np.float64 else lib.oracle_sselect_layout
    fn.restype = ctypes.c_int
    m = fn(ctypes.c_char(howmny.encode()), _ptr(sel), ctypes.c_int(n), _ptr(np.ascontiguousarray(alphai)),
           _ptr(kind), _ptr(idx))
    return kind[:m].copy(), idx[:m].copy()


def normalize(V, kind=None):
    """LAPACK xTGEVC normalisation in place: largest |Re|+|Im| component of each vector = 1."""
    lib = _load(ORACLE_SO)
    n, m = V.shape
    assert V.flags.f_contiguous
    if np.iscomplexobj(V):
        fn = lib.oracle_znormalize if V.dtype == np.complex128 else lib.oracle_cnormalize
        fn(ctypes.c_int(n), ctypes.c_int(m), _ptr(V), ctypes.c_int(max(n, 1)))
    else:
        k = np.ascontiguousarray(kind, dtype=np.int32)
        fn = lib.oracle_dnormalize if V.dtype == np.float64 else lib.oracle_snormalize
        fn(ctypes.c_int(n), ctypes.c_int(m), _ptr(k), _ptr(V), ctypes.c_int(max(n, 1)))
    return V


# ---------------------------------------------------------------------------------------
# residual metrics
# ---------------------------------------------------------------------------------------
def complex_columns(V, kind):
    """Real-interface output -> one complex vector per eigenvalue column (pairs combined:
    column k = v_re + i v_im, column k+1 = v_re - i v_im)."""
    n, m = V.shape
    out = np.zeros((n, m), dtype=np.complex128)
    c = 0
    while c < m:
        if kind[c] == 1:
            z = V[:, c].astype(np.complex128) + 1j * V[:, c + 1]
            out[:, c] = z
            out[:, c + 1] = np.conj(z)
            c += 2
        else:
            out[:, c] = V[:, c]
            c += 1
    return out


def tgsna_e(S, P, VL, VR, kind):
    """Eigenvalue condition numbers, restating LAPACK 3.x dtgsna (JOB = 'E', the reciprocal
    condition number S(ks) of the eigenvalue branch): for a real eigenvalue with left / right
    vectors u, v, s = dlapy2(u^T S v, u^T P v) / (|u| |v|); for a complex pair with columns
    (u0, u1), (v0, v1): uhav = u0'S v0 + u1'S v1, uhavi = u0'S v1 - u1'S v0 (the same with P),
    s = dlapy2(dlapy2(uhav, uhavi), dlapy2(uhbv, uhbvi)) / (dlapy2(|v0|, |v1|) dlapy2(|u0|, |u1|)),
    stored in both columns.  kind: column kinds from real_layout."""
    n, m = VR.shape
    out = np.zeros(m)
    c = 0
    while c < m:
        if kind[c] == 1:
            u0, u1, v0, v1 = VL[:, c], VL[:, c + 1], VR[:, c], VR[:, c + 1]
            sa0, sa1, pb0, pb1 = S @ v0, S @ v1, P @ v0, P @ v1
            av, avi = u0 @ sa0 + u1 @ sa1, u0 @ sa1 - u1 @ sa0
            bv, bvi = u0 @ pb0 + u1 @ pb1, u0 @ pb1 - u1 @ pb0
            rn = np.hypot(np.linalg.norm(v0), np.linalg.norm(v1))
            ln = np.hypot(np.linalg.norm(u0), np.linalg.norm(u1))
            out[c] = out[c + 1] = np.hypot(np.hypot(av, avi), np.hypot(bv, bvi)) / (rn * ln)
            c += 2
        else:
            u, v = VL[:, c], VR[:, c]
            out[c] = np.hypot(u @ (S @ v), u @ (P @ v)) / (np.linalg.norm(u) * np.linalg.norm(v))
            c += 1
    return out


def tgsna_e_complex(S, P, VL, VR):
    """LAPACK 3.x ztgsna (JOB = 'E') restated: s = dlapy2(|u^H S v|, |u^H P v|) / (|u| |v|)
    per column."""
    out = np.zeros(VR.shape[1])
    for c in range(VR.shape[1]):
        u, v = VL[:, c], VR[:, c]
        out[c] = np.hypot(abs(np.vdot(u, S @ v)), abs(np.vdot(u, P @ v))) / (np.linalg.norm(u) * np.linalg.norm(v))
    return out


def eig_alpha_beta(alpha, alphai, beta, idx):
    al = np.asarray(alpha, dtype=np.float64)[idx] + 1j * (np.zeros(len(idx)) if alphai is None
                                                          else np.asarray(alphai, dtype=np.float64)[idx])
    be = np.asarray(beta)[idx]
    return al.astype(np.complex128), be.astype(np.complex128)


def residual_A(S, P, V, al, be, left=False):
    """Metric A per vector: ||(b S - a P) v||_2 / ((|b| ||S||_F + |a| ||P||_F) ||v||_2);
    left vectors use v^H (b S - a P) (tests/test_dtgevc3.cpp:29-223)."""
    S = np.asarray(S, dtype=np.complex128)
    P = np.asarray(P, dtype=np.complex128)
    V = np.asarray(V, dtype=np.complex128)
    nS = np.linalg.norm(S)
    nP = np.linalg.norm(P)
    if left:
        R = (np.conj(V).T @ S).T * be[None, :] - (np.conj(V).T @ P).T * al[None, :]
    else:
        R = (S @ V) * be[None, :] - (P @ V) * al[None, :]
    num = np.linalg.norm(R, axis=0)
    den = (np.abs(be) * nS + np.abs(al) * nP) * np.linalg.norm(V, axis=0)
    den = np.maximum(den, np.finfo(np.float64).tiny)
    return num / den


def residual_B(S, P, V, al, be, left=False, eps=None):
    """Metric B: DGET52-style, 1-norms, result in units of eps (tests/test_lapack_d.cpp:38-168)."""
    S = np.asarray(S, dtype=np.complex128)
    P = np.asarray(P, dtype=np.complex128)
    V = np.asarray(V, dtype=np.complex128)
    if eps is None:
        eps = np.finfo(np.float64).eps
    n = S.shape[0]
    mask = np.triu(np.ones((n, n)), -1)
    nS = np.abs(S * mask).sum(axis=0).max() if n else 1.0
    nP = np.abs(P * mask).sum(axis=0).max() if n else 1.0
    nS = nS if nS else 1.0
    nP = nP if nP else 1.0
    if left:
        R = (np.conj(V).T @ S).T * be[None, :] - (np.conj(V).T @ P).T * al[None, :]
    else:
        R = (S @ V) * be[None, :] - (P @ V) * al[None, :]
    nR = np.abs(R.real).sum(axis=0) + np.abs(R.imag).sum(axis=0)
    nV = np.abs(V.real).sum(axis=0) + np.abs(V.imag).sum(axis=0)
    nV = np.where(nV == 0, 1.0, nV)
    den = (np.abs(be) * nS + (np.abs(al.real) + np.abs(al.imag)) * nP) * nV
    den = np.maximum(den, np.finfo(np.float64).tiny)
    return nR / den / eps


# ---------------------------------------------------------------------------------------
# eigenvector condition numbers for the parity tolerance 100 * eps * cond_k
# ---------------------------------------------------------------------------------------
def eigvec_cond(S, P, alpha, beta, kind, idx, left=False):
    """cond_k = ||(S,P)||_F * (|beta_k| + |alpha_k|) / sigma_min(T_k), with T_k the part of
    beta_k S - alpha_k P that the vector's back (right) or forward (left) substitution
    inverts: rows/cols [0, k) for right vectors, [k + w, n) for left vectors (w = 2 for a
    pair).  A perturbation of the pencil of relative size eps moves the unit-normalised
    vector by about eps * cond_k.  Dense SVDs: keep n small (<= ~200)."""
    S = np.asarray(S, dtype=np.complex128)
    P = np.asarray(P, dtype=np.complex128)
    n = S.shape[0]
    nrm = np.sqrt(np.linalg.norm(S) ** 2 + np.linalg.norm(P) ** 2)
    out = np.zeros(len(idx))
    for c, k in enumerate(idx):
        w = 2 if kind[c] in (1, 2) else 1
        k0 = k if kind[c] != 2 else k - 1
        al, be = alpha[c], beta[c]
        C = be * S - al * P
        T = C[k0 + w:, k0 + w:] if left else C[:k0, :k0]
        if T.shape[0] == 0:
            out[c] = 1.0
            continue
        smin = np.linalg.svd(T, compute_uv=False)[-1]
        out[c] = np.inf if smin == 0 else nrm * (abs(be) + abs(al)) / smin
    return out


def _openblas():
    """The OpenBLAS/LAPACK that ships in this image's SciPy wheel (Fortran names carry a
    scipy_ prefix): an independent LAPACK for cross-checks (xTGEVC)."""
    import glob
    import scipy

    d = os.path.realpath(os.path.join(os.path.dirname(scipy.__file__), "..", "scipy.libs"))
    return _load(glob.glob(os.path.join(d, "libscipy_openblas*.so"))[0])


# ---------------------------------------------------------------------------------------
# FP64 implicit-differentiation reference gradient of selected right eigenvectors
# ---------------------------------------------------------------------------------------
def implicit_grad_right(S, P, idx, V, G):
    """dL/dS, dL/dP for L = sum_c <G[:, c], V[:, c]>, V[:, c] the LAPACK-normalised right
    eigenvector of the real eigenvalue alpha = S[k,k], beta = P[k,k] (k = idx[c]).

    Implicit function theorem on (beta S - alpha P)[0:k, 0:k+1] [x; 1] = 0 with x the raw
    vector (x_k = 1), v = x / max|x|: with g'' = g - (g.v) sign(v_p) e_p restricted to rows
    < k and lambda = T^{-T} g'' (T = (beta S - alpha P)[0:k, 0:k], dense solve):
        dL/dS[:k, :k+1] = -beta lambda v^T,  dL/dP[:k, :k+1] = alpha lambda v^T,
        dL/dS[k,k] += lambda^T (P v)[:k],  dL/dP[k,k] += -lambda^T (S v)[:k],
    kept on the entries the forward reads (S: upper Hessenberg band, P: upper triangle; zero
    below).  Dense O(k^3) per column: small n only."""
    n = S.shape[0]
    dS = np.zeros((n, n))
    dP = np.zeros((n, n))
    for c, k in enumerate(idx):
        v = V[:, c]
        g = G[:, c]
        al, be = S[k, k], P[k, k]
        p = int(np.argmax(np.abs(v)))
        gg = g.copy()
        gg[p] -= np.dot(g, v) * np.sign(v[p])
        if k == 0:
            continue
        T = (be * S - al * P)[:k, :k]
        lam = np.linalg.solve(T.T, gg[:k])
        dS[:k, :k + 1] += -be * np.outer(lam, v[:k + 1])
        dP[:k, :k + 1] += al * np.outer(lam, v[:k + 1])
        dS[k, k] += lam @ (P[:k, :] @ v)
        dP[k, k] += -(lam @ (S[:k, :] @ v))
    # the forward reads only the upper Hessenberg band of S and the upper triangle of P: entries
    # below carry no gradient
    dS *= np.triu(np.ones((n, n)), -1)
    dP *= np.triu(np.ones((n, n)))
    return dS, dP
