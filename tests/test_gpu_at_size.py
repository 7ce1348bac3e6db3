"""The benchmark configurations at their own sizes, through size-independent properties:
every output finite, m / info exact, the LAPACK normalisation, and the backward-error
residual (metric A) of EVERY vector computed on the device (xtgevc3_amd.residuals).

* n = 4096, side R, howmny B, the quasi-triangular N(0,1) construction, FP64 and split-int8;
  and its graded variant (rows / columns scaled by 10^U(-120,120)), whose residuals are held to
  those of the oracle (the reference algorithm) column by column;
* ztgevc3 n = 8192, side B, howmny B (complex triangular, unitary Q and Z), FP64 and split-int8;
* n = 16384 with eigenvalues clustered in 64 groups (relative spread 1e-6), side R, howmny A
  and B, FP64 and split-int8, and the FP32 interface stgevc3 on the FP32-rounded pencil.

Bounds: FP64 10 n eps; split-int8 max(10 n eps, 2.5e-11), its documented floor of 40-bit digit
products; FP32 interface 10 n eps_32."""
import numpy as np
import pytest

from oracle import oracle as O
from xtgevc3_amd import pencils as G
from xtgevc3_amd.residuals import real_column_kinds, residuals_complex, residuals_real

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def _bound(prec, n):
    return 10 * n * EPS if prec == "fp64" else max(10 * n * EPS, 2.5e-11)


def _orth(n, seed, complex_=False):
    """Random orthogonal / unitary matrix on the device (Q of a Gaussian QR), as a tensor holding
    its column-major storage."""
    import torch

    g = torch.Generator(device="cuda").manual_seed(seed)
    A = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
    if complex_:
        A = torch.complex(A, torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g))
    Qm, R = torch.linalg.qr(A)
    d = torch.diagonal(R)
    Qm = Qm * (d / d.abs())[None, :]
    return Qm.t().contiguous()


def _normalised_real(V_t, kinds):
    """max(|Re| + |Im|) of every eigenvector column (pairs over both columns), on the device."""
    import torch

    V = V_t[: len(kinds)]
    k = torch.as_tensor(kinds, device=V.device)
    a = V.abs()
    nxt = torch.roll(a, -1, dims=0)
    s = torch.where((k == 1)[:, None], a + nxt, a)
    colmax = s.max(dim=1).values
    first = k != 2
    return colmax[first].cpu().numpy()


def _prec(X, name):
    return X.PREC_FP64 if name == "fp64" else X.PREC_SPLIT_I8


@pytest.mark.parametrize("prec", ["fp64", "split_i8"])
def test_n4096_right_backtransformed_every_residual(prec):
    import torch

    import xtgevc3_amd as X

    n = 4096
    S, P, ar, ai, be = G.real_quasi_triangular(n, 1)
    dS, dP = X.fortran_device(S), X.fortran_device(P)
    dZ = _orth(n, 2)
    V = dZ.clone()
    m, info = X.dtgevc3_device("R", "B", None, dS, dP, ar, ai, be, VR=V, precision=_prec(X, prec))
    torch.cuda.synchronize()
    assert (m, info) == (n, 0)
    assert bool(torch.isfinite(V).all())
    kinds, idx = real_column_kinds(ai, howmny="B")
    assert np.allclose(_normalised_real(V, kinds), 1.0, rtol=0, atol=64 * EPS)
    res = residuals_real(dS, dP, V, ar, ai, be, kinds, idx, Z_t=dZ)
    print(f"n={n} R/B {prec}: max residual {res.max() / (n * EPS):.3g} n eps over {m} vectors")
    assert res.max() <= _bound(prec, n), f"{res.max() / (n * EPS):.3g} n eps"


@pytest.mark.timeout(300)
def test_n4096_graded_residuals_no_worse_than_the_reference_algorithm():
    """Graded pencil (10^+-120 row / column scaling) at n = 4096, side R, howmny A: outputs
    finite, m / info as the oracle's; columns that come out zero or denormal (complex pairs
    whose 2x2 block lies far below the pencil norm: shifted entries below the double range)
    only where the oracle's columns degenerate too; and every other column's residual at most
    max(10 n eps, 2 x the oracle's residual for that column)."""
    import torch

    import xtgevc3_amd as X

    n = 4096
    S, P, ar, ai, be = G.real_quasi_triangular(n, 5, grade=120.0)
    dS, dP = X.fortran_device(S), X.fortran_device(P)
    V = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    m, info = X.dtgevc3_device("R", "A", None, dS, dP, ar, ai, be, VR=V)
    torch.cuda.synchronize()
    oR = np.zeros((n, n), order="F")
    om, oinfo, _ = O.tgevc3("d", "R", "A", None, S, P, ar, be, alphai=ai, VR=oR)
    assert (m, info) == (om, oinfo) == (n, 0)
    assert bool(torch.isfinite(V).all())
    kinds, idx = real_column_kinds(ai, howmny="A")
    mg = _normalised_real(V, kinds)
    kind_o, _ = O.real_layout("A", None, ai)
    oc = O.complex_columns(oR, kind_o)
    mo = np.max(np.abs(oc.real) + np.abs(oc.imag), axis=0)[kind_o != 2]
    dg, do = ~(mg > 1e-300), ~(mo > 1e-300)
    assert not np.any(dg & ~do)
    rg = residuals_real(dS, dP, V, ar, ai, be, kinds, idx)
    with np.errstate(all="ignore"):
        O.normalize(oR, kind_o)
    ro = residuals_real(dS, dP, X.fortran_device(np.where(np.isfinite(oR), oR, 0.0)), ar, ai, be, kinds, idx)
    first = kinds != 2
    keep = ~dg & ~do
    rgf, rof = rg[first][keep], ro[first][keep]
    worse = rgf > np.maximum(10 * n * EPS, 2.0 * rof)
    print(f"graded n={n}: {dg.sum()} degenerate GPU columns ({do.sum()} oracle); residual <= 10 n eps in "
          f"{(rgf <= 10 * n * EPS).sum()} of {keep.sum()} eigenvalues (oracle {(rof <= 10 * n * EPS).sum()})")
    assert not np.any(worse), f"{worse.sum()} columns worse than the reference algorithm"


@pytest.mark.timeout(300)
@pytest.mark.parametrize("prec", ["fp64", "split_i8"])
def test_ztgevc3_n8192_both_sides_every_residual(prec):
    import torch

    import xtgevc3_amd as X

    n = 8192
    S, P, al, be = G.complex_triangular(n, 3)
    dS, dP = X.fortran_device(S), X.fortran_device(P)
    del S, P
    dQ, dZ = _orth(n, 4, complex_=True), _orth(n, 5, complex_=True)
    VL, VR = dQ.clone(), dZ.clone()
    m, info = X.ztgevc3_device("B", "B", None, dS, dP, al, be, VL=VL, VR=VR, precision=_prec(X, prec))
    torch.cuda.synchronize()
    assert (m, info) == (n, 0)
    idx = np.arange(n)
    for name, V, W, left in (("R", VR, dZ, False), ("L", VL, dQ, True)):
        assert bool(torch.isfinite(torch.view_as_real(V)).all()), name
        nrm = (V.real.abs() + V.imag.abs()).max(dim=1).values.cpu().numpy()
        assert np.allclose(nrm, 1.0, rtol=0, atol=64 * EPS), name
        res = residuals_complex(dS, dP, V, al, be, idx, Z_t=W, left=left)
        print(f"ztgevc3 n={n} B/B {prec} {name}: max residual {res.max() / (n * EPS):.3g} n eps")
        assert res.max() <= _bound(prec, n), f"{name}: {res.max() / (n * EPS):.3g} n eps"


@pytest.fixture(scope="module")
def clustered_16384():
    import xtgevc3_amd as X

    n = 16384
    S, P, ar, ai, be = G.real_quasi_triangular(n, 0, cluster_groups=64)
    return n, S, P, ar, ai, be, X.fortran_device(S), X.fortran_device(P)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("howmny", ["A", "B"])
def test_n16384_clustered_both_forms_every_residual(clustered_16384, howmny):
    import torch

    import xtgevc3_amd as X

    n, S, P, ar, ai, be, dS, dP = clustered_16384
    dZ = _orth(n, 7) if howmny == "B" else None
    kinds, idx = real_column_kinds(ai, howmny=howmny)
    outs = {}
    for prec in ("fp64", "split_i8"):
        V = dZ.clone() if howmny == "B" else torch.zeros((n, n), dtype=torch.float64, device="cuda")
        m, info = X.dtgevc3_device("R", howmny, None, dS, dP, ar, ai, be, VR=V, precision=_prec(X, prec))
        torch.cuda.synchronize()
        assert (m, info) == (n, 0)
        assert bool(torch.isfinite(V).all())
        assert np.allclose(_normalised_real(V, kinds), 1.0, rtol=0, atol=64 * EPS)
        res = residuals_real(dS, dP, V, ar, ai, be, kinds, idx, Z_t=dZ)
        print(f"clustered n={n} R/{howmny} {prec}: max residual {res.max() / (n * EPS):.3g} n eps")
        assert res.max() <= _bound(prec, n), f"{prec}: {res.max() / (n * EPS):.3g} n eps"
        outs[prec] = V
    print(f"  max |V_split - V_fp64| = {float((outs['split_i8'] - outs['fp64']).abs().max()):.3g}")


@pytest.mark.timeout(300)
@pytest.mark.parametrize("howmny", ["A", "B"])
def test_stgevc3_n16384_clustered_every_residual(clustered_16384, howmny):
    """The FP32 interface (host pointers, the drop-in call) on the FP32-rounded clustered pencil
    at n = 16384: computed in FP64 on the device and rounded to FP32; residual of every vector in
    FP64 against the FP32 pencil within 10 n eps_32."""
    import torch

    import xtgevc3_amd as X

    n, S, P, ar, ai, be, _, _ = clustered_16384
    f = np.float32
    S32, P32 = np.asfortranarray(S.astype(f)), np.asfortranarray(P.astype(f))
    ar32, ai32, be32 = ar.astype(f), ai.astype(f), be.astype(f)
    if howmny == "B":
        dZ = _orth(n, 8)
        Z32 = dZ.to(torch.float32)
        VR = np.asfortranarray(Z32.t().cpu().numpy())
        dZ32 = Z32.to(torch.float64)
    else:
        VR = np.zeros((n, n), dtype=f, order="F")
        dZ32 = None
    m, info, _ = X.stgevc3("R", howmny, None, S32, P32, ar32, ai32, be32, VR=VR)
    assert (m, info) == (n, 0)
    assert np.all(np.isfinite(VR))
    kinds, idx = real_column_kinds(ai32, howmny=howmny)
    dV = X.fortran_device(VR.astype(np.float64))
    assert np.allclose(_normalised_real(dV, kinds), 1.0, rtol=0, atol=64 * np.finfo(f).eps)
    res = residuals_real(X.fortran_device(S32.astype(np.float64)), X.fortran_device(P32.astype(np.float64)), dV,
                         ar32.astype(np.float64), ai32.astype(np.float64), be32.astype(np.float64), kinds, idx,
                         Z_t=dZ32)
    e32 = np.finfo(f).eps
    print(f"stgevc3 clustered n={n} R/{howmny}: max residual {res.max() / (n * e32):.3g} n eps_32")
    assert res.max() <= 10 * n * e32
