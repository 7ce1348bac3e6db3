"""Pins xtgevc3_amd.residuals (the device residual used by bench.py and the at-size GPU tests)
against the oracle's metric A (oracle.residual_A, the formula of tests/test_dtgevc3.cpp:29-223)
on the CPU, on vectors perturbed by 1e-6 so that the residual is far above rounding level:
right and left vectors, with and without back-transform, real pairs and complex pencils."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from xtgevc3_amd import pencils as G
from xtgevc3_amd.residuals import real_column_kinds, residuals_complex, residuals_real


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).T))


@pytest.mark.parametrize("howmny", ["A", "B", "S"])
def test_real_residuals_match_metric_a(howmny):
    n = 60
    rng = np.random.default_rng(1)
    S, P, ar, ai, be = G.real_quasi_triangular(n, 3, offdiag=0.2)
    Z, Q = G.random_orthogonal(n, 4), G.random_orthogonal(n, 5)
    sel = (rng.random(n) < 0.4).astype(np.int32) if howmny == "S" else None
    back = howmny == "B"
    VL = Q.copy(order="F") if back else np.zeros((n, n), order="F")
    VR = Z.copy(order="F") if back else np.zeros((n, n), order="F")
    m, info, _ = O.tgevc3("d", "B", howmny, sel, S, P, ar, be, alphai=ai, VL=VL, VR=VR)
    kind, idx = O.real_layout(howmny, sel, ai)
    k2, i2 = real_column_kinds(ai, select=sel, howmny=howmny)
    assert np.array_equal(kind, k2) and np.array_equal(idx, i2)
    al, bb = O.eig_alpha_beta(ar, ai, be, idx)
    A, B = (Q @ S @ Z.T, Q @ P @ Z.T) if back else (S, P)
    for left, V, W in ((False, VR, Z), (True, VL, Q)):
        V = np.asfortranarray(V[:, :m] + 1e-6 * rng.standard_normal((n, m)))
        r1 = O.residual_A(A, B, O.complex_columns(V, kind), al, bb, left=left)
        r2 = residuals_real(_t(S), _t(P), _t(V), ar, ai, be, kind, idx, Z_t=_t(W) if back else None, left=left)
        assert np.allclose(r2, r1, rtol=1e-4, atol=0)


def test_complex_residuals_match_metric_a():
    n = 60
    rng = np.random.default_rng(2)
    S, P, al, be = G.complex_triangular(n, 6, offdiag=0.2)
    Z, Q = G.random_unitary(n, 7), G.random_unitary(n, 8)
    VL, VR = Q.copy(order="F"), Z.copy(order="F")
    O.tgevc3("z", "B", "B", None, S, P, al, be, VL=VL, VR=VR)
    A, B = Q @ S @ Z.conj().T, Q @ P @ Z.conj().T
    for left, V, W in ((False, VR, Z), (True, VL, Q)):
        V = np.asfortranarray(V + 1e-6 * (rng.standard_normal(V.shape) + 1j * rng.standard_normal(V.shape)))
        r1 = O.residual_A(A, B, V, al, be, left=left)
        r2 = residuals_complex(_t(S), _t(P), _t(V), al, be, np.arange(n), Z_t=_t(W), left=left)
        assert np.allclose(r2, r1, rtol=1e-4, atol=0)
