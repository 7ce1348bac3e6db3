"""torch.library operators over the device-pointer C ABI, with autograd through the
hand-written backward kernel.

    V = torch.ops.xtgevc3.right_eigvecs(S, P, alphar, alphai, beta, select, precision)

S, P: (n, n) float64 CUDA tensors in ordinary (row-major) torch layout holding the real
quasi-triangular S and upper-triangular P; alphar, alphai, beta: (n,) float64 CPU tensors (the
reference's interface takes the eigenvalues as inputs, dtgevc3.cpp:1150-1160); select: (n,)
int32 CPU tensor, the reference's select array (howmny='S', selection rule of
dtgevc3.cpp:1192-1211); precision: xtgevc3_amd.PREC_FP64, PREC_SPLIT_I8 or PREC_MIXED_F32.
Returns V (n, m): column c is the normalised right eigenvector of the c-th selected
eigenvalue, a complex pair as its real and imaginary part in two columns.  Gradients exist for
real, finite selected eigenvalues; alpha_k = S_kk and beta_k = P_kk are differentiated as
functions of the diagonals.  There is no CPU kernel: the operators are registered for CUDA
(HIP) tensors only.
"""
import numpy as np
import torch

import xtgevc3_amd as X


def _host(t, dtype):
    return np.ascontiguousarray(t.detach().to("cpu", dtype).numpy())


@torch.library.custom_op("xtgevc3::right_eigvecs", mutates_args=(), device_types="cuda")
def right_eigvecs(S: torch.Tensor, P: torch.Tensor, alphar: torch.Tensor, alphai: torch.Tensor,
                  beta: torch.Tensor, select: torch.Tensor, precision: int = 0) -> torch.Tensor:
    n = S.shape[0]
    sel = _host(select, torch.int32)
    ar, ai, be = (_host(t, torch.float64) for t in (alphar, alphai, beta))
    # column-major storage of S == row-major storage of S^T
    dS = S.detach().t().contiguous()
    dP = P.detach().t().contiguous()
    # each selected index contributes at most two columns
    mm = max(1, min(n, 2 * int(np.count_nonzero(sel))))
    Vt = torch.zeros((mm, n), dtype=torch.float64, device=S.device)
    m, info = X.dtgevc3_device("R", "S", sel, dS, dP, ar, ai, be, VR=Vt, mm=mm, precision=precision)
    if info != 0:
        raise RuntimeError(f"xtgevc3::right_eigvecs: info = {info}")
    return Vt[:m].t().contiguous()


@right_eigvecs.register_fake
def _(S, P, alphar, alphai, beta, select, precision=0):
    m = torch.library.get_ctx().new_dynamic_size()
    return S.new_empty((S.shape[0], m))


@torch.library.custom_op("xtgevc3::right_eigvecs_backward", mutates_args=(), device_types="cuda")
def right_eigvecs_backward(S: torch.Tensor, P: torch.Tensor, alphar: torch.Tensor, alphai: torch.Tensor,
                           beta: torch.Tensor, select: torch.Tensor, V: torch.Tensor, gV: torch.Tensor,
                           precision: int = 0) -> tuple[torch.Tensor, torch.Tensor]:
    sel = _host(select, torch.int32)
    ar, ai, be = (_host(t, torch.float64) for t in (alphar, alphai, beta))
    dS = S.detach().t().contiguous()
    dP = P.detach().t().contiguous()
    Vt = V.detach().t().contiguous()   # column-major storage of V
    G = gV.detach().t().contiguous()   # column-major storage of dL/dV
    gS, gP = X.dtgevc3_backward_device(dS, dP, ar, ai, be, sel, Vt, G, nbatch=1, precision=precision)
    # column-major gradient storage -> torch layout
    return gS.t().contiguous(), gP.t().contiguous()


@right_eigvecs_backward.register_fake
def _(S, P, alphar, alphai, beta, select, V, gV, precision=0):
    return S.new_empty(S.shape), P.new_empty(P.shape)


def _setup_context(ctx, inputs, output):
    S, P, alphar, alphai, beta, select, precision = inputs
    ctx.save_for_backward(S, P, alphar, alphai, beta, select, output)
    ctx.precision = precision


def _backward(ctx, gV):
    S, P, alphar, alphai, beta, select, V = ctx.saved_tensors
    gS, gP = torch.ops.xtgevc3.right_eigvecs_backward(S, P, alphar, alphai, beta, select, V, gV.contiguous(),
                                                      ctx.precision)
    return gS, gP, None, None, None, None, None


right_eigvecs.register_autograd(_backward, setup_context=_setup_context)
