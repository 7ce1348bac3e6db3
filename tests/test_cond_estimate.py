"""Pins tests/helpers.eigvec_cond_sampled (the cond_k of the sampled-column parity tests)
against its definition evaluated with a dense SVD per column, on the columns where the SVD
itself is accurate (cond_k < 1e12; beyond that 100 eps cond_k exceeds 1 and the tolerance
is vacuous)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import eigvec_cond_sampled
from xtgevc3_amd import pencils as G


def _dense(S, P, alpha, beta, kind, idx, left):
    S = np.asarray(S, dtype=np.complex128)
    P = np.asarray(P, dtype=np.complex128)
    fs, fp = 1.0 / np.abs(S).max(), 1.0 / np.abs(P).max()
    S, P = S * fs, P * fp
    nS, nP = np.linalg.norm(S), np.linalg.norm(P)
    out = []
    for c, k in enumerate(idx):
        w = 2 if kind[c] in (1, 2) else 1
        k0 = k if kind[c] != 2 else k - 1
        al, be = alpha[c] * fs, beta[c] * fp
        sl = slice(k0 + w, None) if left else slice(0, k0)
        T = be * S[sl, sl] - al * P[sl, sl]
        out.append(1.0 if T.shape[0] == 0 else (abs(be) * nS + abs(al) * nP) / np.linalg.svd(T, compute_uv=False)[-1])
    return np.array(out)


@pytest.mark.parametrize("left", [False, True])
@pytest.mark.parametrize("case", ["real", "real_scaled", "complex"])
def test_sampled_cond_matches_dense_svd(case, left):
    if case == "complex":
        S, P, al0, be0 = G.complex_triangular(90, 4)
        idx = np.arange(90)
        kind = np.zeros(90, dtype=np.int32)
        al, bb = al0, be0
    else:
        S, P, ar, ai, be = G.real_quasi_triangular(120, 6, offdiag=0.05)
        if case == "real_scaled":
            # separate scalings of S and P (with alpha, beta) leave cond_k unchanged
            S, P, ar, ai, be = S * 1e200, P * 1e-150, ar * 1e200, ai * 1e200, be * 1e-150
        kind, idx = O.real_layout("A", None, ai)
        al, bb = O.eig_alpha_beta(ar, ai, be, idx)
    est = eigvec_cond_sampled(S, P, al, bb, kind, idx, left=left)
    Su, Pu = (S, P) if case != "real_scaled" else (S * 1e-200, P * 1e150)
    alu, bbu = (al, bb) if case != "real_scaled" else (al * 1e-200, bb * 1e150)
    ref = _dense(Su, Pu, alu, bbu, kind, idx, left)
    ok = ref < 1e12
    assert ok.sum() >= 20
    assert np.allclose(est[ok], ref[ok], rtol=1e-4)
