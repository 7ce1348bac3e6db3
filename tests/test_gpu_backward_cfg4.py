"""The differentiable configuration at its own size: dtgevc3 side R, howmny S with 256 selected
real eigenvalues at n = 2048, loss sum_k <w_k, x_k> over 32 cotangent batches, the hand-written
backward in one call for all batches.

* On the construction itself (quasi-triangular, N(0,1) off-diagonals) the eigenvectors are
  exponentially ill-conditioned in n (triangular-solve growth: cond_k grows like c^k), so the
  exact gradients of many selected columns exceed the double range.  Checked: the 32-batch
  call equals single-batch calls entry for entry (same non-finite pattern), info counts the
  overflowed adjoints (and is 0 exactly when every entry is finite), and gradient entries
  that do not depend on an overflowed adjoint stay finite.
* On the same construction with off-diagonals x 0.003 and the 256 best-separated real
  eigenvalues (the gradients carry digits): FP64 central finite differences (h = 1e-6, 200 random
  entries of S and P inside their patterns, alpha_k = S_kk and beta_k = P_kk moved with the
  diagonal) and the FP64 implicit-differentiation gradient of the oracle (dense solves on the
  CPU), for one batch each, plus the batched-vs-single agreement."""
import numpy as np
import pytest

from oracle import oracle as O
from xtgevc3_amd import pencils as G

pytestmark = pytest.mark.gpu

N, NSEL, NBATCH = 2048, 256, 32


def _selection(ar, ai, be, mode, seed=0):
    n = ar.shape[0]
    real_k = np.array([k for k in range(n) if ai[k] == 0.0 and (k == 0 or ai[k - 1] == 0.0)])
    if mode == "random":
        pick = np.sort(np.random.default_rng(seed).choice(real_k, size=NSEL, replace=False))
    else:
        lam = (ar + 1j * ai) / be
        gap = np.array([np.partition(np.abs(lam - lam[k]), 1)[1] for k in real_k])
        pick = np.sort(real_k[np.argsort(-gap)[:NSEL]])
    sel = np.zeros(n, dtype=np.int32)
    sel[pick] = 1
    return sel, pick


def _forward(X, dS, dP, ar, ai, be, sel):
    import torch

    Vt = torch.zeros((NSEL, N), dtype=torch.float64, device="cuda")
    m, info = X.dtgevc3_device("R", "S", sel, dS, dP, ar, ai, be, VR=Vt, mm=NSEL)
    assert (m, info) == (NSEL, 0)
    return Vt


def _single_batches_agree(X, dS, dP, ar, ai, be, sel, Vt, Gt, gS, gP, batches):
    import torch

    for b in batches:
        Gb = Gt[b * NSEL:(b + 1) * NSEL].contiguous()
        s1, p1, _ = X.dtgevc3_backward_device(dS, dP, ar, ai, be, sel, Vt, Gb, nbatch=1, return_info=True)
        for one, many in ((s1, gS[b * N:(b + 1) * N]), (p1, gP[b * N:(b + 1) * N])):
            fin1, finm = torch.isfinite(one), torch.isfinite(many)
            assert torch.equal(fin1, finm), f"batch {b}: non-finite patterns differ"
            a, c = one[fin1], many[finm]
            scale = float(c.abs().max()) if c.numel() else 0.0
            assert float((a - c).abs().max()) <= 1e-12 * max(scale, 1e-300), f"batch {b}"


@pytest.mark.timeout(300)
def test_cfg4_nominal_pencil_batched_equals_single_and_overflow_is_reported():
    import torch

    import xtgevc3_amd as X

    S, P, ar, ai, be = G.real_quasi_triangular(N, 0)
    sel, pick = _selection(ar, ai, be, "random")
    dS, dP = X.fortran_device(S), X.fortran_device(P)
    Vt = _forward(X, dS, dP, ar, ai, be, sel)
    gen = torch.Generator(device="cuda").manual_seed(1)
    Gt = torch.randn((NBATCH * NSEL, N), dtype=torch.float64, device="cuda", generator=gen)
    with _nowarn():
        gS, gP, info = X.dtgevc3_backward_device(dS, dP, ar, ai, be, sel, Vt, Gt, nbatch=NBATCH, return_info=True)
    torch.cuda.synchronize()
    fin = float(torch.isfinite(gS).sum() + torch.isfinite(gP).sum()) / (2.0 * gS.numel())
    print(f"nominal pencil: info = {info} overflowed adjoints of {NBATCH * NSEL}; finite gradient entries {fin:.4f}")
    assert 0 <= info <= NBATCH * NSEL
    assert (info == 0) == bool(torch.isfinite(gS).all() and torch.isfinite(gP).all())
    # an entry (j, i) of dS_b only involves columns c with j <= k_c and i < k_c (v_c is zero
    # below row k_c, lambda_c from row k_c on), plus the diagonal terms at (k_c, k_c): the last
    # column of every batch depends on no selected column and is exactly zero, overflow or not
    if pick.max() < N - 1:
        for b in range(NBATCH):
            assert float(gS[b * N + N - 1].abs().max()) == 0.0  # tensor row = matrix column
    _single_batches_agree(X, dS, dP, ar, ai, be, sel, Vt, Gt, gS, gP, (0, 13, NBATCH - 1))


class _nowarn:
    """Silences the backward's overflow warning (expected on the nominal pencil)."""

    def __enter__(self):
        import warnings

        self._cm = warnings.catch_warnings()
        self._cm.__enter__()
        warnings.simplefilter("ignore", RuntimeWarning)
        return self

    def __exit__(self, *a):
        return self._cm.__exit__(*a)


@pytest.mark.timeout(300)
def test_cfg4_scaled_pencil_fd_implicit_and_batches():
    import torch

    import xtgevc3_amd as X

    S, P, ar, ai, be = G.real_quasi_triangular(N, 0, offdiag=0.003)
    sel, pick = _selection(ar, ai, be, "separated")
    dS, dP = X.fortran_device(S), X.fortran_device(P)
    Vt = _forward(X, dS, dP, ar, ai, be, sel)
    gen = torch.Generator(device="cuda").manual_seed(2)
    Gt = torch.randn((NBATCH * NSEL, N), dtype=torch.float64, device="cuda", generator=gen)
    gS, gP, info = X.dtgevc3_backward_device(dS, dP, ar, ai, be, sel, Vt, Gt, nbatch=NBATCH, return_info=True)
    torch.cuda.synchronize()
    assert info == 0
    assert bool(torch.isfinite(gS).all() and torch.isfinite(gP).all())
    _single_batches_agree(X, dS, dP, ar, ai, be, sel, Vt, Gt, gS, gP, (0, NBATCH - 1))

    # implicit-differentiation reference (oracle, dense FP64 solves) for batch 0
    V = Vt.cpu().numpy().T.copy()
    G0 = Gt[:NSEL].cpu().numpy().T.copy()
    kind, idx = O.real_layout("S", sel, ai)
    iS, iP = O.implicit_grad_right(S, P, idx, V, G0)
    g0S = gS[:N].cpu().numpy().T
    g0P = gP[:N].cpu().numpy().T
    scale = max(np.abs(iS).max(), np.abs(iP).max())
    err = max(np.abs(g0S - iS).max(), np.abs(g0P - iP).max()) / scale
    print(f"scaled pencil: max |g - g_implicit| / max |g| = {err:.3g} (max |g| = {scale:.3g})")
    assert err <= 1e-9

    # central differences on 200 entries for batch 1: L = <G_1, V>
    G1 = Gt[NSEL:2 * NSEL]
    g1S = gS[N:2 * N].cpu().numpy().T
    g1P = gP[N:2 * N].cpu().numpy().T
    in_pair = np.zeros(N, dtype=bool)
    for k in range(N - 1):
        if ai[k] != 0.0 and (k == 0 or not in_pair[k]):
            in_pair[k] = in_pair[k + 1] = True
    rng = np.random.default_rng(3)
    picks = []
    while len(picks) < 200:
        j, i = sorted(rng.integers(0, N, 2))
        if in_pair[j] or in_pair[i]:
            continue  # the pairs' (alpha, beta) are functions of their 2x2 blocks
        picks.append((int(rng.integers(0, 2)), j, i))
    h = 1e-6

    def loss(arv, bev):
        Vx = _forward(X, dS, dP, arv, ai, bev, sel)
        return float((Vx * G1).sum())

    worst = 0.0
    for which, j, i in picks:
        mat = dS if which == 0 else dP
        orig = float(mat[i, j])
        vals = []
        for sgn in (1.0, -1.0):
            mat[i, j] = orig + sgn * h
            arv, bev = ar.copy(), be.copy()
            if i == j:
                (arv if which == 0 else bev)[j] = orig + sgn * h
            vals.append(loss(arv, bev))
        mat[i, j] = orig
        fd = (vals[0] - vals[1]) / (2 * h)
        g = (g1S if which == 0 else g1P)[j, i]
        worst = max(worst, abs(fd - g) / (1e-4 * max(1.0, abs(g))))
        assert abs(fd - g) <= 1e-4 * max(1.0, abs(g)), (which, j, i, fd, g)
    print(f"finite differences: worst |fd - g| / (1e-4 max(1, |g|)) = {worst:.3g}")
