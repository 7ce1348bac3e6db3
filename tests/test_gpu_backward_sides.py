"""Backward of left eigenvectors and of back-transformed vectors (xtgevc3_dtgevc3_backward_full_device):
the adjoint of the reference's left sweep (src/dtgevc3.cpp:1476-1663) and of the back-transform
VR := Z X, VL := Q Y (:1443-1454, :1635-1657), for selected real eigenvalues.

Checked against FP64 central finite differences of the loss L = sum_b <G_b, V> recomputed with
the forward on the GPU (h = 1e-6, entries of S and P inside their patterns with alpha_k = S_kk and
beta_k = P_kk moved with the diagonal, and entries of Z / Q), against the oracle's
implicit-differentiation gradient (left vectors, through the reversed transposed pencil), and for
consistency: batched equals single-batch, and side R without Z equals the right-only entry."""
import numpy as np
import pytest

from oracle import oracle as O
from xtgevc3_amd import pencils as G

pytestmark = pytest.mark.gpu

N = 48


def _pencil(seed=3):
    S, P, ar, ai, be = G.real_quasi_triangular(N, seed, offdiag=0.2)
    reals = [k for k in range(N) if ai[k] == 0.0 and (k == 0 or ai[k - 1] == 0.0)]
    rng = np.random.default_rng(seed)
    pick = np.sort(rng.choice(reals, size=8, replace=False))
    sel = np.zeros(N, dtype=np.int32)
    sel[pick] = 1
    in_pair = np.zeros(N, dtype=bool)
    k = 0
    while k < N:
        if ai[k] != 0.0 and k + 1 < N:
            in_pair[k] = in_pair[k + 1] = True
            k += 2
        else:
            k += 1
    return S, P, ar, ai, be, sel, pick, in_pair


def _forward(X, side, dS, dP, ar, ai, be, sel, pick, dZ=None):
    """Selected vectors (n x m tensor holding column-major storage) of side `side`; with dZ the
    back-transformed ones, taken from the howmny='B' output at the selected eigenvalues' columns."""
    import torch

    m = len(pick)
    if dZ is None:
        V = torch.zeros((m, N), dtype=torch.float64, device="cuda")
        kw = {"VL": V} if side == "L" else {"VR": V}
        mm, info = X.dtgevc3_device(side, "S", sel, dS, dP, ar, ai, be, mm=m, **kw)
        assert (mm, info) == (m, 0)
        return V
    V = dZ.clone()
    kw = {"VL": V} if side == "L" else {"VR": V}
    mm, info = X.dtgevc3_device(side, "B", None, dS, dP, ar, ai, be, **kw)
    assert (mm, info) == (N, 0)
    kind, idx = O.real_layout("B", None, ai)
    cols = [int(np.nonzero((idx == k) & (kind == 0))[0][0]) for k in pick]
    return V[cols].contiguous()


def _fd_check(X, side, S, P, ar, ai, be, sel, pick, in_pair, Gt, gS, gP, Z=None, gZ=None, count=60):
    import torch

    dS, dP = X.fortran_device(S), X.fortran_device(P)
    dZ = X.fortran_device(Z) if Z is not None else None
    rng = np.random.default_rng(11)
    h = 1e-6

    def loss(arv, bev):
        Vx = _forward(X, side, dS, dP, arv, ai, bev, sel, pick, dZ)
        return float((Vx * Gt).sum())

    picks = []
    while len(picks) < count:
        which = int(rng.integers(0, 3 if Z is not None else 2))
        if which < 2:
            j, i = sorted(rng.integers(0, N, 2))
            if in_pair[j] or in_pair[i]:
                continue  # the pairs' (alpha, beta) are functions of their 2x2 blocks
        else:
            j, i = (int(x) for x in rng.integers(0, N, 2))
        picks.append((which, j, i))
    worst = 0.0
    for which, j, i in picks:
        mat = (dS, dP, dZ)[which]
        orig = float(mat[i, j])
        vals = []
        for sg in (1.0, -1.0):
            mat[i, j] = orig + sg * h
            arv, bev = ar.copy(), be.copy()
            if which < 2 and i == j:
                (arv if which == 0 else bev)[j] = orig + sg * h
            vals.append(loss(arv, bev))
        mat[i, j] = orig
        fd = (vals[0] - vals[1]) / (2 * h)
        g = (gS, gP, gZ)[which][j, i]
        worst = max(worst, abs(fd - g) / (1e-5 * max(1.0, abs(g))))
        assert abs(fd - g) <= 1e-5 * max(1.0, abs(g)), (("S", "P", "Z")[which], j, i, fd, g)
    return worst


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("back", [False, True])
def test_backward_sides_match_finite_differences(side, back):
    import torch

    import xtgevc3_amd as X

    if side == "R" and not back:
        pytest.skip("covered by the right-only entry's tests (and by test_full_entry_right_equals_right_entry)")
    S, P, ar, ai, be, sel, pick, in_pair = _pencil()
    Z = G.random_orthogonal(N, 5) if back else None
    dS, dP = X.fortran_device(S), X.fortran_device(P)
    dZ = X.fortran_device(Z) if back else None
    V = _forward(X, side, dS, dP, ar, ai, be, sel, pick, dZ)
    m = len(pick)
    gen = torch.Generator(device="cuda").manual_seed(4)
    Gt = torch.randn((m, N), dtype=torch.float64, device="cuda", generator=gen)
    out = X.dtgevc3_backward_full_device(side, dS, dP, ar, ai, be, sel, V, Gt, nbatch=1, Z=dZ)
    torch.cuda.synchronize()
    gS, gP = out[0].cpu().numpy().T, out[1].cpu().numpy().T
    gZ = out[2].cpu().numpy().T if back else None
    assert np.all(np.isfinite(gS)) and np.all(np.isfinite(gP))
    worst = _fd_check(X, side, S, P, ar, ai, be, sel, pick, in_pair, Gt, gS, gP, Z, gZ)
    print(f"side {side} back={back}: worst |fd - g| / (1e-5 max(1, |g|)) = {worst:.3g}")


def test_left_backward_matches_implicit_gradient_of_the_working_pencil():
    """dL/dS, dL/dP of left vectors against the oracle's implicit gradient of the right vectors
    of the reversed transposed pencil J S^T J, J P^T J, mapped back with J (.)^T J."""
    import torch

    import xtgevc3_amd as X

    S, P, ar, ai, be, sel, pick, _ = _pencil(7)
    dS, dP = X.fortran_device(S), X.fortran_device(P)
    V = _forward(X, "L", dS, dP, ar, ai, be, sel, pick)
    m = len(pick)
    Gt = torch.randn((m, N), dtype=torch.float64, device="cuda", generator=torch.Generator(device="cuda").manual_seed(9))
    gS, gP = X.dtgevc3_backward_full_device("L", dS, dP, ar, ai, be, sel, V, Gt, nbatch=1)
    J = np.eye(N)[::-1]
    Sw, Pw = J @ S.T @ J, J @ P.T @ J
    idxw = np.sort(N - 1 - pick)
    Vl = V.cpu().numpy().T  # n x m, columns in ascending eigenvalue order
    Gl = Gt.cpu().numpy().T
    Vw = (J @ Vl)[:, ::-1]
    Gw = (J @ Gl)[:, ::-1]
    iSw, iPw = O.implicit_grad_right(Sw, Pw, idxw, np.ascontiguousarray(Vw), np.ascontiguousarray(Gw))
    iS, iP = J @ iSw.T @ J, J @ iPw.T @ J
    g0S, g0P = gS.cpu().numpy().T, gP.cpu().numpy().T
    scale = max(np.abs(iS).max(), np.abs(iP).max())
    err = max(np.abs(g0S - iS).max(), np.abs(g0P - iP).max()) / scale
    print(f"left vs implicit: {err:.3g} (max |g| {scale:.3g})")
    assert err <= 1e-10


@pytest.mark.parametrize("side", ["L", "R"])
def test_full_entry_batches_and_right_equals_right_entry(side):
    """Three batches in one call equal three single-batch calls (with the back-transform), and
    side R without Z equals xtgevc3_dtgevc3_backward_device bit for bit."""
    import torch

    import xtgevc3_amd as X

    S, P, ar, ai, be, sel, pick, _ = _pencil(13)
    dS, dP = X.fortran_device(S), X.fortran_device(P)
    Z = G.random_orthogonal(N, 6)
    dZ = X.fortran_device(Z)
    V = _forward(X, side, dS, dP, ar, ai, be, sel, pick, dZ)
    m = len(pick)
    Gt = torch.randn((3 * m, N), dtype=torch.float64, device="cuda", generator=torch.Generator(device="cuda").manual_seed(2))
    aS, aP, aZ = X.dtgevc3_backward_full_device(side, dS, dP, ar, ai, be, sel, V, Gt, nbatch=3, Z=dZ)
    for b in range(3):
        sS, sP, sZ = X.dtgevc3_backward_full_device(side, dS, dP, ar, ai, be, sel, V, Gt[b * m:(b + 1) * m].contiguous(),
                                                     nbatch=1, Z=dZ)
        for one, many in ((sS, aS), (sP, aP), (sZ, aZ)):
            ref = many[b * N:(b + 1) * N]
            assert float((one - ref).abs().max()) <= 1e-13 * max(float(ref.abs().max()), 1.0)
    if side == "R":
        Vr = _forward(X, "R", dS, dP, ar, ai, be, sel, pick)
        G1 = Gt[:m].contiguous()
        fS, fP = X.dtgevc3_backward_full_device("R", dS, dP, ar, ai, be, sel, Vr, G1, nbatch=1)
        rS, rP = X.dtgevc3_backward_device(dS, dP, ar, ai, be, sel, Vr, G1, nbatch=1)
        assert torch.equal(fS, rS) and torch.equal(fP, rP)
