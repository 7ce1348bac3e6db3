"""Backward-error residuals of computed eigenvectors, every column, on the GPU.

Metric A of the reference's tests (tests/test_dtgevc3.cpp:29-223, tests/test_ztgevc3.cpp:29-124),
per vector v with eigenvalue (alpha, beta):

    right:  ||(beta S - alpha P) x||_2 / ((|beta| ||S||_F + |alpha| ||P||_F) ||x||_2)
    left:   ||x^H (beta S - alpha P)||_2 / (same), computed as ||beta S^T conj(x) - alpha P^T conj(x)||

For back-transformed vectors (howmny = 'B', V = Z X or Q X with Z, Q orthogonal / unitary) the
residual of V against the pencil (Q S Z^H, Q P Z^H) equals that of x = Z^H v (right) or
x = Q^H v (left) against (S, P), so x is formed first and the pencil stays triangular.

Arguments are torch tensors holding column-major storage (the layout of the C ABI: a tensor
T of shape (cols, ld) holds the matrix T.T), as the ``*_device`` entries take them.  The
products run in FP64 (complex128) through torch on the device; this is a check of results,
not part of the eigenvector path.
"""
import numpy as np


def real_column_kinds(alphai, m_cols=None, select=None, howmny="A"):
    """Column kinds of the real interface's output (0 real, 1 real part of a pair, 2 its
    imaginary part) and the eigenvalue index of each column, by the reference's selection
    rule (src/dtgevc3.cpp:1192-1207: a pair starts at k when alphai[k] != 0 and k + 1 < n)."""
    ai = np.asarray(alphai)
    n = ai.shape[0]
    kinds, idx = [], []
    k = 0
    while k < n:
        pair = ai[k] != 0.0 and k + 1 < n
        w = 2 if pair else 1
        sel = howmny in ("A", "B") or (select is not None and (select[k] != 0 or (pair and select[k + 1] != 0)))
        if sel:
            if pair:
                kinds += [1, 2]
                idx += [k, k + 1]
            else:
                kinds.append(0)
                idx.append(k)
        k += w
    return np.array(kinds, dtype=np.int32), np.array(idx, dtype=np.int64)


def _xcols(V_t, m, B_t, complex_):
    import torch

    V = V_t[:m].t()  # n x m view of the column-major output
    if B_t is None:
        return V.contiguous()
    Bh = B_t.conj() if complex_ else B_t  # B_t holds B^T: B^H v = conj(B_t) v
    return torch.matmul(Bh.contiguous(), V.contiguous())


def residuals_real(S_t, P_t, V_t, alphar, alphai, beta, kinds, idx, Z_t=None, left=False):
    """Metric A of every column of a real-interface output V (m = len(kinds) columns), pairs
    combined into complex vectors (both columns of a pair get the pair's value).  Z_t: the
    back-transform matrix (Z for right, Q for left vectors) when V was back-transformed.
    Returns a numpy array of m residuals."""
    import torch

    m = len(kinds)
    X = _xcols(V_t, m, Z_t, False)  # n x m
    S = S_t.t()
    P = P_t.t()
    dev = X.device
    kinds_t = torch.as_tensor(kinds, device=dev)
    first = torch.nonzero(kinds_t != 2).flatten()  # one column per eigenvalue
    im_col = torch.where(kinds_t[first] == 1, first + 1, first)
    has_im = (kinds_t[first] == 1).to(X.dtype)
    Xc = torch.complex(X[:, first], X[:, im_col] * has_im[None, :])
    if left:
        # x^H (b S - a P) = conj((b S - a P)^T conj(x))^T: use S^T = S_t
        Sx = torch.matmul(S_t.to(torch.complex128), Xc.conj())
        Px = torch.matmul(P_t.to(torch.complex128), Xc.conj())
    else:
        Sx = torch.matmul(S.contiguous().to(torch.complex128), Xc)
        Px = torch.matmul(P.contiguous().to(torch.complex128), Xc)
    idx_first = np.asarray(idx)[first.cpu().numpy()]
    al = torch.as_tensor(np.asarray(alphar, dtype=np.float64)[idx_first] +
                         1j * np.asarray(alphai, dtype=np.float64)[idx_first], device=dev)
    be = torch.as_tensor(np.asarray(beta, dtype=np.float64)[idx_first].astype(np.complex128), device=dev)
    R = Sx * be[None, :] - Px * al[None, :]
    nS = torch.linalg.norm(S_t)
    nP = torch.linalg.norm(P_t)
    den = (be.abs() * nS + al.abs() * nP) * torch.linalg.norm(Xc, dim=0)
    r = (torch.linalg.norm(R, dim=0) / torch.clamp(den, min=np.finfo(np.float64).tiny)).cpu().numpy()
    out = np.zeros(m)
    fc = first.cpu().numpy()
    out[fc] = r
    pair = np.asarray(kinds)[fc] == 1
    out[fc[pair] + 1] = r[pair]
    return out


def residuals_complex(S_t, P_t, V_t, alpha, beta, idx, Z_t=None, left=False):
    """Metric A of every column of a complex-interface output V (m = len(idx) columns, column c
    the eigenvector of eigenvalue idx[c]); Z_t as in residuals_real.  Numpy array of m values."""
    import torch

    m = len(idx)
    X = _xcols(V_t, m, Z_t, True)
    dev = X.device
    if left:
        Sx = torch.matmul(S_t.contiguous(), X.conj())
        Px = torch.matmul(P_t.contiguous(), X.conj())
    else:
        Sx = torch.matmul(S_t.t().contiguous(), X)
        Px = torch.matmul(P_t.t().contiguous(), X)
    al = torch.as_tensor(np.asarray(alpha, dtype=np.complex128)[idx], device=dev)
    be = torch.as_tensor(np.asarray(beta, dtype=np.complex128)[idx], device=dev)
    R = Sx * be[None, :] - Px * al[None, :]
    den = (be.abs() * torch.linalg.norm(S_t) + al.abs() * torch.linalg.norm(P_t)) * torch.linalg.norm(X, dim=0)
    return (torch.linalg.norm(R, dim=0) / torch.clamp(den, min=np.finfo(np.float64).tiny)).cpu().numpy()
