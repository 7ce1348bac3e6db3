"""torch.autograd wrapper of the selected right eigenvectors (dtgevc3, side='R',
howmny='S' or 'A', no back-transform), for testing the hand-written backward kernel.

    V = RightEigvecs.apply(S, P, alphar, alphai, beta, select, precision)

S, P: (n, n) float64 CUDA tensors in ordinary (row-major) torch layout; alphar/alphai/beta:
host numpy arrays (the reference's interface takes the eigenvalues as inputs).  For the
selected real eigenvalues the backward differentiates alpha_k = S[k, k] and
beta_k = P[k, k] as functions of the diagonals; dL/dS is nonzero only on the upper Hessenberg
band and dL/dP only on the upper triangle (the forward reads nothing below them).  Returns
V (n, m): column c is the
normalised right eigenvector of the c-th selected eigenvalue (LAPACK convention).
precision (xtgevc3_amd.PREC_*) selects the form of the forward's and the backward's far
updates: FP64, split-int8 or FP32 products.
"""
import numpy as np
import torch

import xtgevc3_amd as X


class RightEigvecs(torch.autograd.Function):
    @staticmethod
    def forward(ctx, S, P, alphar, alphai, beta, select, precision=X.PREC_FP64):
        n = S.shape[0]
        # column-major storage of S == row-major storage of S^T
        dS = S.detach().t().contiguous()
        dP = P.detach().t().contiguous()
        howmny = "A" if select is None else "S"
        m_cols = n if select is None else int(np.count_nonzero(select))
        Vt = torch.zeros((max(m_cols, 1), n), dtype=torch.float64, device=S.device)
        m, info = X.dtgevc3_device("R", howmny, select, dS, dP, alphar, alphai, beta, VR=Vt, mm=Vt.shape[0],
                                   precision=precision)
        if info != 0:
            raise RuntimeError(f"dtgevc3 forward info={info}")
        Vt = Vt[:m]
        ctx.save_for_backward(dS, dP, Vt)
        ctx.ev = (alphar, alphai, beta, select)
        ctx.precision = precision
        return Vt.t()

    @staticmethod
    def backward(ctx, gV):
        dS_cm, dP_cm, Vt = ctx.saved_tensors
        alphar, alphai, beta, select = ctx.ev
        n = dS_cm.shape[0]
        G = gV.t().contiguous()  # column-major storage of dL/dV
        gS, gP = X.dtgevc3_backward_device(dS_cm, dP_cm, alphar, alphai, beta, select, Vt, G, nbatch=1,
                                           precision=ctx.precision)
        # column-major dS storage -> torch layout
        return gS.t(), gP.t(), None, None, None, None, None


def right_eigvecs(S, P, alphar, alphai, beta, select=None, precision=X.PREC_FP64):
    return RightEigvecs.apply(S, P, alphar, alphai, beta, select, precision)
