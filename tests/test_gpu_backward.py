"""GPU backward (adjoint) kernel: gradients of selected right eigenvectors w.r.t. S and P,
checked against FP64 central finite differences of the GPU forward and against the FP64
implicit-differentiation reference on the CPU (oracle.implicit_grad_right)."""
import numpy as np
import pytest

from oracle import oracle as O
from xtgevc3_amd import pencils as G

pytestmark = pytest.mark.gpu


def _problem(n, nsel, seed, offdiag=0.2):
    # scaled off-diagonals keep the selected eigenvectors well conditioned (cond <= ~1e8), so
    # FP64 gradients carry meaningful digits; N(0,1) pencils reach cond ~1e18 by n = 48
    S, P, ar, ai, be = G.real_quasi_triangular(n, seed, offdiag=offdiag)
    rng = np.random.default_rng(seed)
    real_k = np.array([k for k in range(n) if ai[k] == 0.0 and (k == 0 or ai[k - 1] == 0.0)])
    pick = np.sort(rng.choice(real_k, size=min(nsel, len(real_k)), replace=False))
    sel = np.zeros(n, dtype=np.int32)
    sel[pick] = 1
    return S, P, ar, ai, be, sel, pick


def _forward_gpu(S, P, ar, ai, be, sel):
    import torch

    from xtgevc3_amd.autograd import right_eigvecs

    St = torch.tensor(S, device="cuda")
    Pt = torch.tensor(P, device="cuda")
    # the eigenvalue inputs follow the (perturbed) diagonals for the selected columns
    ar = ar.copy()
    be = be.copy()
    k = np.nonzero(sel)[0]
    ar[k] = S[k, k]
    be[k] = P[k, k]
    return right_eigvecs(St, Pt, ar, ai, be, sel)


@pytest.mark.parametrize("n,nsel,seed,offdiag", [(24, 6, 1, 0.3), (48, 10, 2, 0.2), (70, 12, 3, 0.15)])
def test_backward_matches_implicit_and_fd(n, nsel, seed, offdiag):
    import torch

    S, P, ar, ai, be, sel, pick = _problem(n, nsel, seed, offdiag)
    rng = np.random.default_rng(seed + 100)
    W = rng.standard_normal((n, len(pick)))
    St = torch.tensor(S, device="cuda", requires_grad=True)
    Pt = torch.tensor(P, device="cuda", requires_grad=True)
    from xtgevc3_amd.autograd import right_eigvecs

    V = right_eigvecs(St, Pt, ar, ai, be, sel)
    L = (V * torch.tensor(W, device="cuda")).sum()
    L.backward()
    gS = St.grad.cpu().numpy()
    gP = Pt.grad.cpu().numpy()
    # implicit-diff reference on the CPU from the oracle's forward
    VR = np.zeros((n, n), order="F")
    O.tgevc3("d", "R", "S", sel, S, P, ar, be, alphai=ai, VR=VR)
    kind, idx = O.real_layout("S", sel, ai)
    Vo = O.normalize(VR[:, :len(pick)].copy(order="F"), kind)
    rS, rP = O.implicit_grad_right(S, P, idx, Vo, W)
    scale = max(np.abs(rS).max(), np.abs(rP).max())
    al, bb = O.eig_alpha_beta(ar, ai, be, idx)
    cond = O.eigvec_cond(S, P, al, bb, kind, idx).max()
    tol = max(1e-9, 10 * np.finfo(float).eps * cond)
    assert np.abs(gS - rS).max() <= tol * scale, (np.abs(gS - rS).max() / scale, cond)
    assert np.abs(gP - rP).max() <= tol * scale, (np.abs(gP - rP).max() / scale, cond)
    # central finite differences (h = 1e-6) of the GPU forward on random support entries
    h = 1e-6
    mask = np.triu(np.ones((n, n))) != 0  # entries that keep the quasi-triangular structure
    cand = np.argwhere(mask)
    pts = cand[rng.choice(len(cand), size=min(60, len(cand)), replace=False)]
    # and entries below the band, which the forward never reads: gradient and quotient are 0
    low = np.argwhere(np.tril(np.ones((n, n)), -2) != 0)
    pts = np.concatenate([pts, low[rng.choice(len(low), size=8, replace=False)]])
    assert np.all(gS[np.tril(np.ones((n, n)), -2) != 0] == 0.0)
    assert np.all(gP[np.tril(np.ones((n, n)), -1) != 0] == 0.0)
    worst = 0.0
    for which, base in (("S", S), ("P", P)):
        for (i, j) in pts:
            if which == "P" and i > j:
                continue
            Lv = []
            for sgn in (1, -1):
                Sx, Px = S.copy(), P.copy()
                (Sx if which == "S" else Px)[i, j] += sgn * h
                Vx = _forward_gpu(Sx, Px, ar, ai, be, sel).cpu().numpy()
                Lv.append(float(np.sum(Vx * W)))
            fd = (Lv[0] - Lv[1]) / (2 * h)
            an = (gS if which == "S" else gP)[i, j]
            worst = max(worst, abs(fd - an) / max(scale, 1e-300))
    assert worst < 1e-6, worst


@pytest.mark.parametrize("n,nb", [(90, 4), (90, 8), (197, 8)])
def test_backward_batched_matches_single(n, nb):
    # nb = 8: the outer products deal whole batches to the XCDs (batch counts divisible by 8);
    # n = 197: odd, partial edge tiles and tiles wholly below the band
    import torch

    import xtgevc3_amd as X

    nsel = 16
    S, P, ar, ai, be, sel, pick = _problem(n, nsel, 9)
    m = len(pick)
    dS = X.fortran_device(S)
    dP = X.fortran_device(P)
    Vt = torch.zeros((m, n), dtype=torch.float64, device="cuda")
    X.dtgevc3_device("R", "S", sel, dS, dP, ar, ai, be, VR=Vt, mm=m)
    rng = np.random.default_rng(5)
    Gs = rng.standard_normal((n, m * nb))
    Gt = torch.tensor(np.ascontiguousarray(Gs.T), device="cuda")
    bS, bP = X.dtgevc3_backward_device(dS, dP, ar, ai, be, sel, Vt, Gt, nbatch=nb)
    assert bool(torch.isfinite(bS).all()) and bool(torch.isfinite(bP).all())
    for b in range(nb):
        # tensor row j holds column j: dS is zero below row j + 1 of column j, dP below row j
        assert not bool(torch.triu(bS[b * n:(b + 1) * n], diagonal=2).any())
        assert not bool(torch.triu(bP[b * n:(b + 1) * n], diagonal=1).any())
    for b in range(nb):
        Gb = torch.tensor(np.ascontiguousarray(Gs[:, b * m:(b + 1) * m].T), device="cuda")
        sS, sP = X.dtgevc3_backward_device(dS, dP, ar, ai, be, sel, Vt, Gb, nbatch=1)
        assert torch.allclose(bS[b * n:(b + 1) * n], sS, rtol=1e-12, atol=1e-12 * float(sS.abs().max()))
        assert torch.allclose(bP[b * n:(b + 1) * n], sP, rtol=1e-12, atol=1e-12 * float(sP.abs().max()))


def test_backward_rejects_complex_selection():
    import torch

    import xtgevc3_amd as X

    n = 30
    S, P, ar, ai, be = G.real_quasi_triangular(n, 4)
    k = int(np.nonzero(ai > 0)[0][0])
    sel = np.zeros(n, dtype=np.int32)
    sel[k] = 1
    dS = X.fortran_device(S)
    dP = X.fortran_device(P)
    Vt = torch.zeros((2, n), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        X.dtgevc3_backward_device(dS, dP, ar, ai, be, sel, Vt, Vt.clone(), nbatch=1)


@pytest.mark.parametrize("prec,tolf", [(3, 100.0), (1, 1e9)])
def test_backward_mixed_matches_fp64(prec, tolf):
    """Mixed forward + mixed backward (far updates of both sweeps as split-int8 digit products,
    or as FP32 products) against the FP64 forward + backward, at a size with far updates
    (n = 400: one group of four 64-row blocks, then far updates), for the 32 best separated
    real eigenvalues.  Bound: tolf * eps * cond_max of the selected eigenvectors, relative to
    the gradient's largest entry."""
    import torch

    import xtgevc3_amd as X

    n, nb, nsel = 400, 3, 32
    S, P, ar, ai, be = G.real_quasi_triangular(n, 12, offdiag=0.02)
    lam = (ar + 1j * ai) / be
    real_k = np.array([k for k in range(n) if ai[k] == 0.0 and (k == 0 or ai[k - 1] == 0.0)])
    gap = np.array([np.min(np.abs(np.delete(lam, k) - lam[k])) for k in real_k])
    pick = np.sort(real_k[np.argsort(-gap)[:nsel]])
    sel = np.zeros(n, dtype=np.int32)
    sel[pick] = 1
    kind, idx = O.real_layout("S", sel, ai)
    al, bb = O.eig_alpha_beta(ar, ai, be, idx)
    cond = O.eigvec_cond(S, P, al, bb, kind, idx).max()
    dS = X.fortran_device(S)
    dP = X.fortran_device(P)
    rng = np.random.default_rng(17)
    Gt = torch.tensor(rng.standard_normal((nb * nsel, n)), device="cuda")
    out = {}
    for p in (X.PREC_FP64, prec):
        Vt = torch.zeros((nsel, n), dtype=torch.float64, device="cuda")
        m, info = X.dtgevc3_device("R", "S", sel, dS, dP, ar, ai, be, VR=Vt, mm=nsel, precision=p)
        assert (m, info) == (nsel, 0)
        gS, gP = X.dtgevc3_backward_device(dS, dP, ar, ai, be, sel, Vt, Gt, nbatch=nb, precision=p)
        out[p] = (Vt.cpu().numpy(), gS.cpu().numpy(), gP.cpu().numpy())
    V64, gS64, gP64 = out[X.PREC_FP64]
    Vm, gSm, gPm = out[prec]
    scale = max(np.abs(gS64).max(), np.abs(gP64).max())
    err = max(np.abs(gSm - gS64).max(), np.abs(gPm - gP64).max()) / scale
    bound = tolf * np.finfo(float).eps * cond
    assert np.isfinite(err) and err <= bound, (err, bound, cond)
    assert np.abs(Vm - V64).max() <= bound, (np.abs(Vm - V64).max(), bound)


def test_backward_rejects_unknown_precision():
    import torch

    import xtgevc3_amd as X

    n = 20
    S, P, ar, ai, be, sel, pick = _problem(n, 3, 5)
    dS = X.fortran_device(S)
    dP = X.fortran_device(P)
    Vt = torch.zeros((len(pick), n), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        X.dtgevc3_backward_device(dS, dP, ar, ai, be, sel, Vt, Vt.clone(), nbatch=1, precision=7)


@pytest.mark.parametrize("precision", ["fp64", "split_i8"])
def test_backward_on_a_side_stream_is_ordered(precision):
    """The backward on a non-default stream: the cotangents are written on it just before the
    call, the gradients read and the buffers reused on it right after, with no device
    synchronisation in between (the adjoint sweep's lookahead products and, in the split form,
    the pencil digits run on other streams ordered by events); equal bit for bit to the
    default-stream call."""
    import torch

    import xtgevc3_amd as X

    prec = {"fp64": X.PREC_FP64, "split_i8": X.PREC_SPLIT_I8}[precision]
    n, nsel, nb = 700, 64, 4
    S, P, ar, ai, be, sel, pick = _problem(n, nsel, 12)
    m = len(pick)
    dS = X.fortran_device(S)
    dP = X.fortran_device(P)
    Vt = torch.zeros((m, n), dtype=torch.float64, device="cuda")
    X.dtgevc3_device("R", "S", sel, dS, dP, ar, ai, be, VR=Vt, mm=m)
    Gh = torch.tensor(np.random.default_rng(8).standard_normal((m * nb, n)), device="cuda")
    refS, refP = X.dtgevc3_backward_device(dS, dP, ar, ai, be, sel, Vt, Gh, nbatch=nb, precision=prec)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        Gt = torch.empty_like(Gh)
        Gt.copy_(Gh)
        gS = torch.empty_like(refS)
        gP = torch.empty_like(refP)
        X.dtgevc3_backward_device(dS, dP, ar, ai, be, sel, Vt, Gt, nbatch=nb, dS=gS, dP=gP, stream=s.cuda_stream,
                                  precision=prec)
        outS = gS.clone()
        outP = gP.clone()
        Gt.fill_(7.0)
        gS.fill_(7.0)
        gP.fill_(7.0)
    torch.cuda.synchronize()
    assert torch.equal(outS, refS) and torch.equal(outP, refP)
