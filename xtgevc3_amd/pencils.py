"""Seeded synthetic generalized-Schur pencils (S, P) with eigenvalues (alpha, beta).

These build the inputs the xTGEVC3 path consumes: S upper (quasi-)triangular, P upper
triangular, and eigenvalues given explicitly, as the reference's interface expects them
(include/tgevc3.h:10-13 takes alphar/alphai/beta as inputs).

* ``real_quasi_triangular`` -- real pencils with 2x2 complex-pair bumps: P diagonal
  |N(0,1)|+0.1, off-diagonal N(0,1); a fraction of the diagonal of S in 2x2 blocks whose
  eigenvalues are complex, with the matching P block diagonal and positive.  Options grade
  the rows/columns by 10^U(-g, g) (overflow scaling), cluster the real eigenvalues in
  groups with relative spread 1e-6 (near-singular shifted blocks), or scale the
  off-diagonal entries (``offdiag``; the eigenvectors of N(0,1) pencils are exponentially
  ill-conditioned in n, a smaller scale gives well-conditioned test problems).
* ``complex_triangular`` -- complex S, P upper triangular, off-diagonal CN(0,1) (times
  ``offdiag``), diag S CN(0,1), diag P |N(0,1)|+0.1 (real).
* ``reference_style_real`` / ``reference_style_complex`` -- the reference tests' own
  construction (tests/test_lapack_d.cpp:170-220, tests/test_lapack_z.cpp:115-140): U(-1,1)
  entries, P diagonal pushed away from zero by +-1, pairs [[a, b], [-b, a]] with P = I.
* ``qz_pencil`` -- A, B ~ N(0,1) reduced by the QZ algorithm (LAPACK xGGES via SciPy):
  (S, P, Q, Z, alpha, beta).
* ``random_orthogonal`` / ``random_unitary`` -- Q factor of the QR of a Gaussian matrix.

All arrays are returned column-major (Fortran order), the layout of the interface.
"""
import numpy as np


def _pair_block(rng):
    """A 2x2 (S, P) block with complex-conjugate generalized eigenvalues and P diagonal > 0.

    det(S - l P) = p1 p2 l^2 - (s11 p2 + s22 p1) l + (s11 s22 - s12 s21); the discriminant
    (s11 p2 - s22 p1)^2 + 4 p1 p2 s12 s21 is made negative by the choice of s21."""
    s11, s22, s12 = rng.standard_normal(3)
    if s12 == 0.0:
        s12 = 1.0
    p1, p2 = np.abs(rng.standard_normal(2)) + 0.1
    d0 = (s11 * p2 - s22 * p1) ** 2
    s21 = -np.sign(s12) * (d0 / (4 * p1 * p2 * abs(s12)) + abs(rng.standard_normal()) + 0.1)
    return s11, s12, s21, s22, p1, p2


def pair_eigenvalue(s11, s12, s21, s22, p1, p2):
    """(alphar, alphai, beta) of the 2x2 block with alphai > 0 and beta = p1 * p2."""
    b = s11 * p2 + s22 * p1
    disc = (s11 * p2 - s22 * p1) ** 2 + 4.0 * p1 * p2 * s12 * s21
    assert disc < 0
    # l = (b +- i sqrt(-disc)) / (2 p1 p2); scale by beta = p1 p2
    return 0.5 * b, 0.5 * np.sqrt(-disc), p1 * p2


def real_quasi_triangular(n, seed, pair_frac=0.4, grade=0.0, cluster_groups=0, cluster_spread=1e-6, offdiag=1.0):
    rng = np.random.default_rng(seed)
    S = np.triu(rng.standard_normal((n, n)), 1) * offdiag
    P = np.triu(rng.standard_normal((n, n)), 1) * offdiag
    ar = np.zeros(n)
    ai = np.zeros(n)
    be = np.zeros(n)
    # choose block starts so that about pair_frac of the diagonal is in 2x2 blocks
    is_pair = np.zeros(n, dtype=bool)
    k = 0
    want = pair_frac / (2.0 - pair_frac)  # probability that a block is a pair
    coin = rng.random(n)
    centres = None
    if cluster_groups:
        centres = rng.standard_normal(cluster_groups)
    while k < n:
        if k + 1 < n and coin[k] < want:
            s11, s12, s21, s22, p1, p2 = _pair_block(rng)
            S[k, k], S[k, k + 1], S[k + 1, k], S[k + 1, k + 1] = s11, s12, s21, s22
            P[k, k], P[k, k + 1], P[k + 1, k + 1] = p1, 0.0, p2
            a_r, a_i, b_ = pair_eigenvalue(s11, s12, s21, s22, p1, p2)
            ar[k] = ar[k + 1] = a_r
            ai[k], ai[k + 1] = a_i, -a_i
            be[k] = be[k + 1] = b_
            is_pair[k] = True
            k += 2
        else:
            p = abs(rng.standard_normal()) + 0.1
            if centres is not None:
                c = centres[rng.integers(cluster_groups)]
                s = p * c * (1.0 + cluster_spread * rng.standard_normal())
            else:
                s = rng.standard_normal()
            S[k, k] = s
            P[k, k] = p
            ar[k], ai[k], be[k] = s, 0.0, p
            k += 1
    if grade:
        dr = 10.0 ** rng.uniform(-grade, grade, n)
        dc = 10.0 ** rng.uniform(-grade, grade, n)
        # keep 2x2 blocks' P diagonal: scale rows/cols of a pair identically
        for k in np.nonzero(is_pair)[0]:
            dr[k + 1] = dr[k]
            dc[k + 1] = dc[k]
        S = dr[:, None] * S * dc[None, :]
        P = dr[:, None] * P * dc[None, :]
        # eigenvalues of the graded pencil: alpha_k / beta_k unchanged.  A pair's block is scaled
        # by one factor f = dr_k dc_k, so its (alpha, beta) from the unscaled block times f
        # stay consistent (recomputing them from the graded entries would overflow)
        scale = dr * dc
        ar = ar * scale
        ai = ai * scale
        be = be * scale
    return (np.asfortranarray(S), np.asfortranarray(P), ar, ai, be)


def complex_triangular(n, seed, offdiag=1.0):
    rng = np.random.default_rng(seed)

    def cn(shape):
        return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2.0)

    S = np.triu(cn((n, n)), 1) * offdiag
    P = np.triu(cn((n, n)), 1) * offdiag
    ds = cn(n)
    dp = np.abs(rng.standard_normal(n)) + 0.1
    S[np.arange(n), np.arange(n)] = ds
    P[np.arange(n), np.arange(n)] = dp
    return np.asfortranarray(S), np.asfortranarray(P), ds.copy(), dp.astype(np.complex128)


def random_orthogonal(n, seed):
    rng = np.random.default_rng(seed)
    q, r = np.linalg.qr(rng.standard_normal((n, n)))
    q *= np.sign(np.diag(r))[None, :]
    return np.asfortranarray(q)


def random_unitary(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    q, r = np.linalg.qr(a)
    d = np.diag(r)
    q *= (d / np.abs(d))[None, :]
    return np.asfortranarray(q)


def reference_style_real(n, seed):
    """The reference tests' generator, with numpy's PCG64 in place of mt19937."""
    rng = np.random.default_rng(seed)
    S = np.triu(rng.uniform(-1, 1, (n, n)))
    P = np.triu(rng.uniform(-1, 1, (n, n)))
    d = np.arange(n)
    P[d, d] += np.where(P[d, d] >= 0, 1.0, -1.0)
    ar = np.zeros(n)
    ai = np.zeros(n)
    be = np.zeros(n)
    k = 0
    while k < n:
        if k < n - 1 and rng.integers(0, 100) < 40:
            P[k, k] = 1.0
            P[k, k + 1] = 0.0
            P[k + 1, k] = 0.0
            P[k + 1, k + 1] = 1.0
            a = S[k, k]
            S[k + 1, k + 1] = a
            b = S[k, k + 1]
            if b == 0.0:
                b = 1.0
            S[k + 1, k] = -b
            ar[k] = ar[k + 1] = a
            ai[k] = abs(b)
            ai[k + 1] = -abs(b)
            be[k] = be[k + 1] = 1.0
            k += 2
        else:
            ar[k] = S[k, k]
            be[k] = P[k, k]
            k += 1
    return np.asfortranarray(S), np.asfortranarray(P), ar, ai, be


def reference_style_complex(n, seed):
    rng = np.random.default_rng(seed)
    S = np.triu(rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n)))
    P = np.triu(rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n)))
    d = np.arange(n)
    P[d, d] += np.where(P[d, d].real >= 0, 1.0, -1.0)
    return np.asfortranarray(S), np.asfortranarray(P), S[d, d].copy(), P[d, d].copy()


def qz_pencil(n, seed, complex_=False):
    """A, B ~ N(0,1) -> generalized Schur form via LAPACK xGGES (SciPy)."""
    import scipy.linalg as sl

    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    B = rng.standard_normal((n, n))
    if complex_:
        A = A + 1j * rng.standard_normal((n, n))
        B = B + 1j * rng.standard_normal((n, n))
        S, P, alpha, beta, Q, Z = sl.ordqz(A, B, sort="lhp", output="complex")
        return (np.asfortranarray(S), np.asfortranarray(P), np.asfortranarray(Q), np.asfortranarray(Z), alpha, beta)
    S, P, alpha, beta, Q, Z = sl.ordqz(A, B, sort="lhp", output="real")
    return (np.asfortranarray(S), np.asfortranarray(P), np.asfortranarray(Q), np.asfortranarray(Z),
            alpha.real.copy(), alpha.imag.copy(), beta.real.copy())


def _diagonal_structure(n, rng, pair_frac, cluster_groups, cluster_spread):
    """The diagonal of ``real_quasi_triangular``: 1x1 entries and 2x2 pair blocks, with their
    eigenvalues.  Returns (entries, ar, ai, be): entries is a list of (row, col, s, p) for S and
    P inside the diagonal blocks."""
    ar, ai, be = np.zeros(n), np.zeros(n), np.zeros(n)
    want = pair_frac / (2.0 - pair_frac)
    coin = rng.random(n)
    centres = rng.standard_normal(cluster_groups) if cluster_groups else None
    rows, cols, sv, pv = [], [], [], []
    k = 0
    while k < n:
        if k + 1 < n and coin[k] < want:
            s11, s12, s21, s22, p1, p2 = _pair_block(rng)
            rows += [k, k, k + 1, k + 1]
            cols += [k, k + 1, k, k + 1]
            sv += [s11, s12, s21, s22]
            pv += [p1, 0.0, 0.0, p2]
            a_r, a_i, b_ = pair_eigenvalue(s11, s12, s21, s22, p1, p2)
            ar[k] = ar[k + 1] = a_r
            ai[k], ai[k + 1] = a_i, -a_i
            be[k] = be[k + 1] = b_
            k += 2
        else:
            p = abs(rng.standard_normal()) + 0.1
            if centres is not None:
                s = p * centres[rng.integers(cluster_groups)] * (1.0 + cluster_spread * rng.standard_normal())
            else:
                s = rng.standard_normal()
            rows.append(k)
            cols.append(k)
            sv.append(s)
            pv.append(p)
            ar[k], be[k] = s, p
            k += 1
    return (np.array(rows), np.array(cols), np.array(sv), np.array(pv)), ar, ai, be


def real_quasi_triangular_device(n, seed, pair_frac=0.4, cluster_groups=0, cluster_spread=1e-6, device="cuda"):
    """The construction of ``real_quasi_triangular`` (without grading) with the off-diagonal
    entries drawn on the device (torch generator), for large n.  Returns (S_t, P_t, ar, ai, be):
    S_t, P_t are device tensors holding the column-major storage of S and P (S = S_t.T), the
    eigenvalues host arrays.  Same distribution as the host generator, other random numbers."""
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    # column-major storage of an upper-triangular matrix = row-major lower-triangular tensor
    S_t = torch.tril(torch.randn(n, n, dtype=torch.float64, device=device, generator=g), -1)
    P_t = torch.tril(torch.randn(n, n, dtype=torch.float64, device=device, generator=g), -1)
    (r, c, sv, pv), ar, ai, be = _diagonal_structure(n, np.random.default_rng(seed), pair_frac, cluster_groups,
                                                     cluster_spread)
    ri = torch.as_tensor(r, device=device)
    ci = torch.as_tensor(c, device=device)
    S_t[ci, ri] = torch.as_tensor(sv, device=device)
    P_t[ci, ri] = torch.as_tensor(pv, device=device)
    return S_t, P_t, ar, ai, be


def complex_triangular_device(n, seed, device="cuda"):
    """``complex_triangular`` with the entries drawn on the device: (S_t, P_t, alpha, beta),
    S_t and P_t device tensors holding column-major storage, alpha and beta host arrays."""
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)

    def cn():
        re = torch.randn(n, n, dtype=torch.float64, device=device, generator=g)
        im = torch.randn(n, n, dtype=torch.float64, device=device, generator=g)
        return torch.complex(re, im) / np.sqrt(2.0)

    S_t = torch.tril(cn(), -1)
    P_t = torch.tril(cn(), -1)
    rng = np.random.default_rng(seed)
    ds = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2.0)
    dp = np.abs(rng.standard_normal(n)) + 0.1
    d = torch.arange(n, device=device)
    S_t[d, d] = torch.as_tensor(ds, device=device)
    P_t[d, d] = torch.as_tensor(dp.astype(np.complex128), device=device)
    return S_t, P_t, ds, dp.astype(np.complex128)
