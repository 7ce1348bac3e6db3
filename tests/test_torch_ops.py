"""torch.library packaging (xtgevc3_amd/torch_ops.py): operator schemas on the CPU; on the GPU,
the operator's values and gradients against the autograd.Function wrapper and the
device-pointer call, and torch.library.opcheck's schema / autograd / fake-tensor checks."""
import numpy as np
import pytest

from xtgevc3_amd import pencils as G


def test_operator_schemas():
    import torch

    import xtgevc3_amd.torch_ops  # noqa: F401  (registers the operators)

    fwd = str(torch.ops.xtgevc3.right_eigvecs.default._schema)
    bwd = str(torch.ops.xtgevc3.right_eigvecs_backward.default._schema)
    assert fwd.startswith("xtgevc3::right_eigvecs(Tensor S, Tensor P, Tensor alphar, Tensor alphai, Tensor beta, "
                          "Tensor select, SymInt precision=0) -> Tensor"), fwd
    assert bwd.endswith("-> (Tensor, Tensor)"), bwd


def _problem(n, nsel, seed, with_pair=False):
    S, P, ar, ai, be = G.real_quasi_triangular(n, seed, offdiag=0.2)
    rng = np.random.default_rng(seed)
    real_k = np.array([k for k in range(n) if ai[k] == 0.0 and (k == 0 or ai[k - 1] == 0.0)])
    sel = np.zeros(n, dtype=np.int32)
    sel[rng.choice(real_k, size=nsel, replace=False)] = 1
    if with_pair:
        sel[int(np.nonzero(ai > 0)[0][0])] = 1
    # the eigenvalue inputs follow the diagonals of the selected real columns
    k = np.nonzero(sel & (ai == 0))[0]
    ar = ar.copy()
    be = be.copy()
    ar[k] = S[k, k]
    be[k] = P[k, k]
    return S, P, ar, ai, be, sel


@pytest.mark.gpu
def test_operator_matches_autograd_function():
    import torch

    import xtgevc3_amd.torch_ops  # noqa: F401
    from xtgevc3_amd.autograd import right_eigvecs

    n = 120
    S, P, ar, ai, be, sel = _problem(n, 9, 31)
    W = torch.tensor(np.random.default_rng(4).standard_normal((n, 9)), device="cuda")
    out = []
    for use_op in (True, False):
        St = torch.tensor(S, device="cuda", requires_grad=True)
        Pt = torch.tensor(P, device="cuda", requires_grad=True)
        if use_op:
            V = torch.ops.xtgevc3.right_eigvecs(St, Pt, torch.tensor(ar), torch.tensor(ai), torch.tensor(be),
                                                torch.tensor(sel), 0)
        else:
            V = right_eigvecs(St, Pt, ar, ai, be, sel)
        (V * W).sum().backward()
        out.append((V.detach().cpu().numpy(), St.grad.cpu().numpy(), Pt.grad.cpu().numpy()))
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_operator_complex_pair_columns():
    import torch

    import xtgevc3_amd as X
    import xtgevc3_amd.torch_ops  # noqa: F401

    n = 90
    S, P, ar, ai, be, sel = _problem(n, 5, 32, with_pair=True)
    V = torch.ops.xtgevc3.right_eigvecs(torch.tensor(S, device="cuda"), torch.tensor(P, device="cuda"),
                                        torch.tensor(ar), torch.tensor(ai), torch.tensor(be), torch.tensor(sel), 0)
    Vt = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    m, info = X.dtgevc3_device("R", "S", sel, X.fortran_device(S), X.fortran_device(P), ar, ai, be, VR=Vt, mm=n)
    assert info == 0 and V.shape == (n, m) and m == 7
    assert torch.equal(V, Vt[:m].t())


@pytest.mark.gpu
def test_operator_opcheck():
    import torch

    import xtgevc3_amd.torch_ops  # noqa: F401

    n = 60
    S, P, ar, ai, be, sel = _problem(n, 4, 33)
    args = (torch.tensor(S, device="cuda", requires_grad=True), torch.tensor(P, device="cuda", requires_grad=True),
            torch.tensor(ar), torch.tensor(ai), torch.tensor(be), torch.tensor(sel), 0)
    torch.library.opcheck(torch.ops.xtgevc3.right_eigvecs.default, args,
                          test_utils=("test_schema", "test_autograd_registration", "test_faketensor"))
