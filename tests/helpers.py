"""Shared checks of the parity tests: run the GPU path and the oracle on the same inputs,
normalise both (LAPACK convention), compare within 100 * eps * cond_k and check residuals."""
import numpy as np

from oracle import oracle as O

EPS = np.finfo(np.float64).eps


def run_real(impl, side, howmny, select, S, P, ar, ai, be, Q=None, Z=None, mm=None):
    """impl: 'gpu' (xtgevc3_amd host entry) or 'oracle'/'ref'. Returns (VL, VR, m, info)."""
    n = S.shape[0]
    cols = n if mm is None else mm
    VL = np.asfortranarray(Q.copy()) if (howmny == "B" and Q is not None) else np.zeros((n, cols), order="F")
    VR = np.asfortranarray(Z.copy()) if (howmny == "B" and Z is not None) else np.zeros((n, cols), order="F")
    if howmny == "B" and Q is None:
        VL = np.asfortranarray(np.eye(n))
    if howmny == "B" and Z is None:
        VR = np.asfortranarray(np.eye(n))
    if impl == "gpu":
        import xtgevc3_amd as X

        m, info, _ = X.dtgevc3(side, howmny, select, S, P, ar, ai, be, VL=VL, VR=VR, mm=mm)
    else:
        m, info, _ = O.tgevc3("d", side, howmny, select, S, P, ar, be, alphai=ai, VL=VL, VR=VR, mm=mm, impl=impl)
    return VL, VR, m, info


def check_real_parity(S, P, ar, ai, be, side="B", howmny="B", select=None, Q=None, Z=None, cond=True,
                      res_bound=10.0, tol_factor=100.0):
    n = S.shape[0]
    gL, gR, gm, ginfo = run_real("gpu", side, howmny, select, S, P, ar, ai, be, Q, Z)
    oL, oR, om, oinfo = run_real("oracle", side, howmny, select, S, P, ar, ai, be, Q, Z)
    assert (gm, ginfo) == (om, oinfo)
    kind, idx = O.real_layout(howmny, select, ai)
    assert len(kind) == gm
    m = gm
    al, bb = O.eig_alpha_beta(ar, ai, be, idx)
    if howmny == "B":
        Qm = np.eye(n) if Q is None else Q
        Zm = np.eye(n) if Z is None else Z
        A = Qm @ S @ Zm.T
        B = Qm @ P @ Zm.T
    else:
        A, B = S, P
    out = {}
    for name, g, o, left in (("R", gR, oR, False), ("L", gL, oL, True)):
        if (name == "R" and side not in "RB") or (name == "L" and side not in "LB"):
            continue
        g = g[:, :m].copy(order="F")
        o = O.normalize(o[:, :m].copy(order="F"), kind)
        assert np.all(np.isfinite(g)), f"{name}: non-finite output"
        # normalisation of the GPU output
        gc = O.complex_columns(g, kind)
        mx = np.max(np.abs(gc.real) + np.abs(gc.imag), axis=0)
        assert np.allclose(mx, 1.0, rtol=0, atol=64 * EPS), f"{name}: not normalised"
        res = O.residual_A(A, B, gc, al, bb, left=left)
        assert res.max() <= res_bound * n * EPS, f"{name}: residual {res.max() / (n * EPS):.3g} n*eps"
        err = np.max(np.abs(g - o), axis=0)
        if cond:
            ck = O.eigvec_cond(S, P, al, bb, kind, idx, left=left)
            tol = tol_factor * EPS * np.maximum(ck, 1.0)
            bad = np.nonzero(err > tol)[0]
            assert bad.size == 0, f"{name}: cols {bad[:8]} err {err[bad[:8]]} tol {tol[bad[:8]]}"
        out[name] = (res.max() / (n * EPS), err.max())
    return out


def run_cplx(impl, side, howmny, select, S, P, al, be, Q=None, Z=None, mm=None, prefix="z"):
    n = S.shape[0]
    dt = np.complex128 if prefix == "z" else np.complex64
    cols = n if mm is None else mm
    if howmny == "B":
        VL = np.asfortranarray((np.eye(n) if Q is None else Q).astype(dt))
        VR = np.asfortranarray((np.eye(n) if Z is None else Z).astype(dt))
    else:
        VL = np.zeros((n, cols), dtype=dt, order="F")
        VR = np.zeros((n, cols), dtype=dt, order="F")
    S = np.asfortranarray(S.astype(dt))
    P = np.asfortranarray(P.astype(dt))
    if impl == "gpu":
        import xtgevc3_amd as X

        m, info, _ = X.tgevc3(prefix, side, howmny, select, S, P, al.astype(dt), be.astype(dt), VL=VL, VR=VR, mm=mm)
    else:
        m, info, _ = O.tgevc3(prefix, side, howmny, select, S, P, al.astype(dt), be.astype(dt), VL=VL, VR=VR, mm=mm,
                              impl=impl)
    return VL, VR, m, info


def check_cplx_parity(S, P, al, be, side="B", howmny="B", select=None, Q=None, Z=None, cond=True, res_bound=10.0,
                      tol_factor=100.0):
    n = S.shape[0]
    gL, gR, gm, ginfo = run_cplx("gpu", side, howmny, select, S, P, al, be, Q, Z)
    oL, oR, om, oinfo = run_cplx("oracle", side, howmny, select, S, P, al, be, Q, Z)
    assert (gm, ginfo) == (om, oinfo)
    m = gm
    idx = np.arange(n) if select is None or howmny != "S" else np.nonzero(select)[0]
    a_k, b_k = al[idx], be[idx]
    if howmny == "B":
        Qm = np.eye(n) if Q is None else Q
        Zm = np.eye(n) if Z is None else Z
        A = Qm @ S @ Zm.conj().T
        B = Qm @ P @ Zm.conj().T
    else:
        A, B = S, P
    kind = np.zeros(m, dtype=np.int32)
    out = {}
    for name, g, o, left in (("R", gR, oR, False), ("L", gL, oL, True)):
        if (name == "R" and side not in "RB") or (name == "L" and side not in "LB"):
            continue
        g = g[:, :m].copy(order="F")
        o = O.normalize(o[:, :m].copy(order="F"))
        assert np.all(np.isfinite(g))
        mx = np.max(np.abs(g.real) + np.abs(g.imag), axis=0)
        assert np.allclose(mx, 1.0, rtol=0, atol=64 * EPS)
        res = O.residual_A(A, B, g, a_k, b_k, left=left)
        assert res.max() <= res_bound * n * EPS, f"{name}: residual {res.max() / (n * EPS):.3g} n*eps"
        err = np.max(np.abs(g - o), axis=0)
        if cond:
            ck = O.eigvec_cond(S, P, a_k, b_k, kind, idx, left=left)
            tol = tol_factor * EPS * np.maximum(ck, 1.0)
            bad = np.nonzero(err > tol)[0]
            assert bad.size == 0, f"{name}: cols {bad[:8]} err {err[bad[:8]]} tol {tol[bad[:8]]}"
        out[name] = (res.max() / (n * EPS), err.max())
    return out


def eigvec_cond_sampled(S, P, alpha, beta, kind, idx, left=False, iters=24):
    """cond_k for the parity tolerance of selected columns at sizes where a dense SVD per
    column is too slow: (|beta_k| ||S||_F + |alpha_k| ||P||_F) / sigma_min(T_k), with T_k the
    part of beta_k S - alpha_k P that the column's substitution inverts (rows / columns [0, k)
    for right vectors, [k + w, n) for left ones, as in oracle.eigvec_cond).  sigma_min comes
    from inverse power iteration on T^H T with one LU of T, which approaches it from above,
    so the estimate never exceeds the true cond_k (the tolerance errs on the strict side).
    S and P are scaled to unit maximum first; the value is invariant under separate scalings
    of S and P (and of alpha and beta with them)."""
    import scipy.linalg as sl

    S = np.asarray(S, dtype=np.complex128)
    P = np.asarray(P, dtype=np.complex128)
    fs = 1.0 / max(np.abs(S).max(), np.finfo(np.float64).tiny)
    fp = 1.0 / max(np.abs(P).max(), np.finfo(np.float64).tiny)
    S = S * fs
    P = P * fp
    nS, nP = np.linalg.norm(S), np.linalg.norm(P)
    rng = np.random.default_rng(0)
    out = np.zeros(len(idx))
    for c, k in enumerate(idx):
        w = 2 if kind[c] in (1, 2) else 1
        k0 = k if kind[c] != 2 else k - 1
        al, be = alpha[c] * fs, beta[c] * fp
        sl_ = slice(k0 + w, None) if left else slice(0, k0)
        T = be * S[sl_, sl_] - al * P[sl_, sl_]
        if T.shape[0] == 0:
            out[c] = 1.0
            continue
        lu = sl.lu_factor(T, check_finite=False)
        if np.any(np.diag(lu[0]) == 0):
            out[c] = np.inf
            continue
        x = rng.standard_normal(T.shape[0]) + 0j
        x /= np.linalg.norm(x)
        est = 0.0
        for _ in range(iters):
            y = sl.lu_solve(lu, x, check_finite=False)          # T^-1 x
            z = sl.lu_solve(lu, y, trans=2, check_finite=False)  # T^-H T^-1 x
            est = np.linalg.norm(z)
            if not np.isfinite(est) or est == 0.0:
                break
            x = z / est
        smin = 1.0 / np.sqrt(est) if est > 0 and np.isfinite(est) else 0.0
        out[c] = np.inf if smin == 0 else (abs(be) * nS + abs(al) * nP) / smin
    return out


def compare_sampled(g, o, cond, tol_factor=100.0):
    """Per-column max |g - o| of normalised vectors against tol_factor * eps * max(cond, 1);
    returns the offending column positions."""
    err = np.max(np.abs(g - o), axis=0)
    tol = tol_factor * EPS * np.maximum(cond, 1.0)
    return np.nonzero(~(err <= tol))[0], err, tol
