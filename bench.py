"""Benchmark of the MI355X xTGEVC3 path on its headline workload.

Workload (one "step"): dtgevc3 side='R', howmny='B' (back-transformed right eigenvectors)
at n = 8192 on a synthetic quasi-triangular pencil -- P upper triangular with diagonal
|N(0,1)|+0.1 and off-diagonal N(0,1), S with off-diagonal N(0,1) and 40 % of its diagonal in
2x2 complex-pair blocks (matching P block diagonal, positive), Z the Q factor of the QR of a
Gaussian matrix.  Inputs are resident in HBM before the timed region; each step restores
VR = Z (device copy, inside the timed region) and calls the device-pointer entry point.

Prints ONE JSON line (see the field list in main()).  --gpus N > 1 exits: this path runs on
one GPU.  The CPU baseline times the reference library compiled from its own sources
(oracle/_ref, linked to the image's OpenBLAS) on the host cores, on the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
PEAK_INT8_TOPS = 5000.0
SPLIT_PRODUCTS = 15  # digit products per multiply-add in the split-int8 form (levels i + j <= 4 of 5 digits)

# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on this workload,
# each scaled by the counter's reading on a plain 1 GiB device copy (FETCH x2.0, WRITE x1.0);
# summary and commands in profiles/dtgevc3_R_B_n8192_hbm_traffic.md
MEASURED_TRAFFIC = {
    ("xtg::k_gemm_f64<1, true, 128, 128, 16, false>", 8192): 11731.3e6,
    ("xtg::k_gemm_f64<0, true, 128, 64, 16, false>", 8192): 409.5e6,
    ("xtg::k_gemm_f64<0, true, 128, 128, 16, false>", 8192): 215.7e6,
    ("xtg::k_gemm_f64<0, true, 64, 64, 32, false>", 8192): 59.2e6,
    ("xtg::k_solve_real<65>", 8192): 40.3e6,
    ("xtg::k_gemm_i8<128, 128, 2, 4> (back-transform)", 8192): 4899.9e6,
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_inputs(n, seed, device, cluster=False, need_z=True, z_host=False):
    import torch

    from xtgevc3_amd import pencils as G

    t0 = time.time()
    # cluster: diagonal ratios s_kk / p_kk drawn from 64 centres with relative spread 1e-6
    S, P, ar, ai, be = G.real_quasi_triangular(n, seed, cluster_groups=64 if cluster else 0)
    dZ = None
    if need_z and z_host:
        # Z from a host QR (no library kernels on the device, e.g. under counter collection)
        dZ = torch.from_numpy(np.ascontiguousarray(G.random_orthogonal(n, seed + 1).T)).to(device)
    elif need_z:
        g = torch.Generator(device=device)
        g.manual_seed(seed + 1)
        A = torch.randn(n, n, dtype=torch.float64, device=device, generator=g)
        Q, R = torch.linalg.qr(A)
        Q = Q * torch.sign(torch.diagonal(R))[None, :]
        # column-major storage of Z == row-major storage of Z^T
        dZ = Q.t().contiguous()
    dS = torch.from_numpy(np.ascontiguousarray(S.T)).to(device)
    dP = torch.from_numpy(np.ascontiguousarray(P.T)).to(device)
    torch.cuda.synchronize()
    log(f"[bench] inputs n={n} ready in {time.time() - t0:.1f}s")
    return S, P, ar, ai, be, dS, dP, dZ


def residual_every_vector(dS, dP, ar, ai, be, dV, howmny="B", dZ=None):
    """Metric A residual of EVERY column of a right-eigenvector output, in units of n*eps,
    computed on the device in FP64 (xtgevc3_amd.residuals): x = Z^T v for back-transformed
    output, r = (beta S - alpha P) x.  Returns (max, number of columns)."""
    from xtgevc3_amd.residuals import real_column_kinds, residuals_real

    n = dS.shape[0]
    kinds, idx = real_column_kinds(ai, howmny=howmny)
    res = residuals_real(dS, dP, dV, ar, ai, be, kinds, idx, Z_t=dZ if howmny == "B" else None)
    return float(res.max()) / (n * np.finfo(np.float64).eps), len(kinds)


def cpu_baseline(n, seed, budget_s=40.0):
    """Reference library (oracle/_ref) on the host cores: the same workload when it fits the
    time budget, else a smaller pencil of the same construction scaled by n^2."""
    from oracle import oracle as O
    from xtgevc3_amd import pencils as G

    if not O.have_ref():
        return None
    cores = len(os.sched_getaffinity(0))
    S, P, ar, ai, be = G.real_quasi_triangular(2048, seed)
    Z = G.random_orthogonal(2048, seed + 1)
    VR = Z.copy(order="F")
    t = time.time()
    O.tgevc3("d", "R", "B", None, S, P, ar, be, alphai=ai, VR=VR, impl="ref")
    t2048 = time.time() - t
    predicted = t2048 * (n / 2048.0) ** 3 * 0.8
    ns = n if predicted <= budget_s else 4096
    if ns != 2048:
        S, P, ar, ai, be = G.real_quasi_triangular(ns, seed)
        Z = G.random_orthogonal(ns, seed + 1)
        VR = Z.copy(order="F")
        t = time.time()
        O.tgevc3("d", "R", "B", None, S, P, ar, be, alphai=ai, VR=VR, impl="ref")
        ts = time.time() - t
    else:
        ts = t2048
    rate = ns / ts
    value = rate * (ns / n) ** 2 if ns != n else rate
    sample = (f"reference dtgevc3 (oracle/_ref, OpenBLAS {cores} threads) side=R howmny=B on the same "
              f"construction at n={ns}: {ts:.2f} s" + ("" if ns == n else f"; eigvecs/s scaled by (n_s/n)^2 to n={n}"))
    return {"value": value, "unit": "eigvecs/s", "cores": cores, "kind": "reference", "sample": sample,
            "seconds": ts, "n_sample": ns}


def _lapack_dtgevc(side, howmny, S, P, VL, VR):
    """LAPACK dtgevc_ of the image's OpenBLAS (the reference tests' comparison point,
    tests/test_lapack_d.cpp:293-297) on column-major arrays, in place.  Returns (m, info)."""
    import ctypes

    from oracle import oracle as O

    n = S.shape[0]
    c = ctypes
    p = lambda a: a.ctypes.data_as(c.c_void_p)
    m, info = c.c_int(0), c.c_int(0)
    work = np.zeros(6 * max(n, 1))
    O._openblas().scipy_dtgevc_(c.byref(c.c_char(side.encode())), c.byref(c.c_char(howmny.encode())), None,
                                c.byref(c.c_int(n)), p(S), c.byref(c.c_int(n)), p(P), c.byref(c.c_int(n)), p(VL),
                                c.byref(c.c_int(VL.shape[0])), p(VR), c.byref(c.c_int(VR.shape[0])),
                                c.byref(c.c_int(n)), c.byref(m), p(work), c.byref(info), c.c_size_t(1), c.c_size_t(1))
    return m.value, info.value


def _timed(fn, reps=3):
    """Median wall time of reps calls (seconds)."""
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return sorted(ts)[len(ts) // 2]


def cpu_legs(n, seed):
    """The CPU comparison points beside the all-cores reference (BASELINE.md plan): the
    reference library single-threaded and LAPACK dtgevc_ (all cores), side R, howmny B on the
    same construction at n_s = 2048, eigvecs/s scaled by (n_s / n)^2 to n (the work grows as
    n^3)."""
    from threadpoolctl import threadpool_limits

    from oracle import oracle as O
    from xtgevc3_amd import pencils as G

    if not O.have_ref():
        return None
    ns = 2048
    S, P, ar, ai, be = G.real_quasi_triangular(ns, seed)
    Z = G.random_orthogonal(ns, seed + 1)
    out = {}
    with threadpool_limits(1):
        t1 = _timed(lambda: O.tgevc3("d", "R", "B", None, S, P, ar, be, alphai=ai, VR=Z.copy(order="F"),
                                     impl="ref"), reps=1)
    out["reference_1_thread"] = {"value": round(ns / t1 * (ns / n) ** 2, 2), "unit": "eigvecs/s", "cores": 1,
                                 "seconds_at_n_sample": round(t1, 3), "n_sample": ns}
    Sf, Pf = np.asfortranarray(S), np.asfortranarray(P)
    tl = _timed(lambda: _lapack_dtgevc("R", "B", Sf, Pf, np.zeros((1, 1), order="F"), Z.copy(order="F")), reps=1)
    out["lapack_dtgevc"] = {"value": round(ns / tl * (ns / n) ** 2, 2), "unit": "eigvecs/s",
                            "cores": len(os.sched_getaffinity(0)), "seconds_at_n_sample": round(tl, 3),
                            "n_sample": ns}
    return out


def config0_cpu(seed):
    """configs[0] on the host cores (the reference's own CPU test path): the reference library
    (all threads and one) and LAPACK dtgevc_ beside it, on the inputs of config0."""
    from threadpoolctl import threadpool_limits

    from oracle import oracle as O
    from xtgevc3_amd import pencils as G

    if not O.have_ref():
        return {}
    S, P, Q, Z, ar, ai, be = G.qz_pencil(512, seed)
    out = {}
    run = lambda: O.tgevc3("d", "B", "B", None, S, P, ar, be, alphai=ai, VL=Q.copy(order="F"), VR=Z.copy(order="F"),
                           impl="ref")
    run()
    out["cpu_reference_ms"] = round(1000 * _timed(run), 3)
    with threadpool_limits(1):
        out["cpu_reference_1_thread_ms"] = round(1000 * _timed(run), 3)
    out["cpu_cores"] = len(os.sched_getaffinity(0))
    Sf, Pf = np.asfortranarray(S), np.asfortranarray(P)
    out["cpu_lapack_dtgevc_ms"] = round(1000 * _timed(lambda: _lapack_dtgevc("B", "B", Sf, Pf, Q.copy(order="F"),
                                                                               Z.copy(order="F"))), 3)
    return out


def config0(seed):
    """dtgevc3 side B, howmny B at n = 512 on the QZ form of an N(0,1) pencil
    (LAPACK dgges via SciPy), VL = Q, VR = Z: the GPU device entry, with the residual of every
    vector (both sides); config0_cpu times the host legs on the same inputs."""
    import torch

    import xtgevc3_amd as X
    from xtgevc3_amd import pencils as G
    from xtgevc3_amd.residuals import real_column_kinds, residuals_real

    n = 512
    S, P, Q, Z, ar, ai, be = G.qz_pencil(n, seed)
    out = {"workload": f"dtgevc3 side=B howmny=B n={n}, QZ form of N(0,1) A, B (dgges), VL=Q, VR=Z"}
    dS, dP = X.fortran_device(S), X.fortran_device(P)
    dQ, dZ = X.fortran_device(Q), X.fortran_device(Z)
    VL, VR = dQ.clone(), dZ.clone()

    def gpu():
        VL.copy_(dQ)
        VR.copy_(dZ)
        X.dtgevc3_device("B", "B", None, dS, dP, ar, ai, be, VL=VL, VR=VR)
        torch.cuda.synchronize()

    gpu()
    t = _timed(gpu, reps=5)
    out["gpu_ms"] = round(1000 * t, 3)
    out["gpu_eigvecs_per_s"] = round(2 * n / t, 1)
    kinds, idx = real_column_kinds(ai, howmny="B")
    eps = np.finfo(np.float64).eps
    out["gpu_max_residual_n_eps"] = {
        side: float("%.4g" % (residuals_real(dS, dP, V, ar, ai, be, kinds, idx, Z_t=W, left=side == "L").max() /
                              (n * eps))) for side, V, W in (("R", VR, dZ), ("L", VL, dQ))}
    return out


def _orth_device(n, seed, complex_=False):
    import torch

    g = torch.Generator(device="cuda").manual_seed(seed)
    A = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
    if complex_:
        A = torch.complex(A, torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g))
    Qm, R = torch.linalg.qr(A)
    d = torch.diagonal(R)
    return (Qm * (d / d.abs())[None, :]).t().contiguous()


def _time_device(fn, steps):
    import torch

    fn()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t) / steps


def config2(seed, steps):
    """ztgevc3 side B, howmny B at n = 8192 (complex triangular pencil, unitary Q
    and Z), plain complex128 and the split-int8 form, with the residual of every vector."""
    import torch

    import xtgevc3_amd as X
    from xtgevc3_amd import pencils as G
    from xtgevc3_amd.residuals import residuals_complex

    n = 8192
    dS, dP, al, be = G.complex_triangular_device(n, seed)
    dQ, dZ = _orth_device(n, seed + 1, True), _orth_device(n, seed + 2, True)
    VL, VR = torch.empty_like(dQ), torch.empty_like(dZ)
    eps = np.finfo(np.float64).eps
    out = {"workload": f"ztgevc3 side=B howmny=B n={n}, complex triangular pencil (CN(0,1)), unitary Q, Z"}
    for name, prec in (("plain", X.PREC_FP64), ("split_int8", X.PREC_SPLIT_I8)):
        def step():
            VL.copy_(dQ)
            VR.copy_(dZ)
            m, info = X.ztgevc3_device("B", "B", None, dS, dP, al, be, VL=VL, VR=VR, precision=prec)
            assert (m, info) == (n, 0)

        t = _time_device(step, steps)
        r = {"ms_per_call": round(1000 * t, 3), "eigvecs_per_s": round(2 * n / t, 1)}
        r["max_residual_n_eps"] = {
            side: float("%.4g" % (residuals_complex(dS, dP, V, al, be, np.arange(n), Z_t=W, left=side == "L").max()
                                  / (n * eps))) for side, V, W in (("R", VR, dZ), ("L", VL, dQ))}
        out[name] = r
    out["split_speedup"] = round(out["plain"]["ms_per_call"] / out["split_int8"]["ms_per_call"], 3)
    del dS, dP, dQ, dZ, VL, VR
    torch.cuda.empty_cache()
    return out


def config3(seed, steps):
    """dtgevc3 side R at n = 16384, eigenvalues clustered in 64 groups (relative
    spread 1e-6), howmny A and B, plain FP64 and split-int8 on the same inputs, plus the FP32
    interface stgevc3 (host pointers, PCIe staging included) on the FP32-rounded pencil; the
    residual of every vector for each."""
    import torch

    import xtgevc3_amd as X
    from xtgevc3_amd import pencils as G
    from xtgevc3_amd.residuals import real_column_kinds, residuals_real

    n = 16384
    dS, dP, ar, ai, be = G.real_quasi_triangular_device(n, seed, cluster_groups=64)
    dZ = _orth_device(n, seed + 3)
    V = torch.empty_like(dZ)
    eps = np.finfo(np.float64).eps
    out = {"workload": f"dtgevc3 side=R n={n}, quasi-triangular pencil (40% pairs), 64 eigenvalue clusters "
                       f"(spread 1e-6), Z orthogonal"}
    for howmny in ("A", "B"):
        kinds, idx = real_column_kinds(ai, howmny=howmny)
        for name, prec in (("plain", X.PREC_FP64), ("split_int8", X.PREC_SPLIT_I8)):
            def step():
                if howmny == "B":
                    V.copy_(dZ)
                m, info = X.dtgevc3_device("R", howmny, None, dS, dP, ar, ai, be, VR=V, precision=prec)
                assert (m, info) == (n, 0)

            t = _time_device(step, steps)
            res = residuals_real(dS, dP, V, ar, ai, be, kinds, idx, Z_t=dZ if howmny == "B" else None)
            out[f"R_{howmny}_{name}"] = {"ms_per_call": round(1000 * t, 3), "eigvecs_per_s": round(n / t, 1),
                                         "max_residual_n_eps": float("%.4g" % (res.max() / (n * eps)))}
        out[f"R_{howmny}_split_speedup"] = round(out[f"R_{howmny}_plain"]["ms_per_call"] /
                                                 out[f"R_{howmny}_split_int8"]["ms_per_call"], 3)
    # FP32 interface through the host-pointer drop-in
    f = np.float32
    S32 = np.asfortranarray(dS.to(torch.float32).t().cpu().numpy())
    P32 = np.asfortranarray(dP.to(torch.float32).t().cpu().numpy())
    Z32d = dZ.to(torch.float32)
    Z32 = np.asfortranarray(Z32d.t().cpu().numpy())
    ar32, ai32, be32 = ar.astype(f), ai.astype(f), be.astype(f)
    VRh = Z32.copy(order="F")

    def host():
        VRh[...] = Z32
        m, info, _ = X.stgevc3("R", "B", None, S32, P32, ar32, ai32, be32, VR=VRh)
        assert (m, info) == (n, 0)

    host()
    t = _timed(host, reps=3)
    kinds, idx = real_column_kinds(ai32, howmny="B")
    res = residuals_real(X.fortran_device(S32.astype(np.float64)), X.fortran_device(P32.astype(np.float64)),
                         X.fortran_device(VRh.astype(np.float64)), ar32.astype(np.float64), ai32.astype(np.float64),
                         be32.astype(np.float64), kinds, idx, Z_t=Z32d.to(torch.float64))
    e32 = np.finfo(f).eps
    out["stgevc3_R_B_host_api"] = {"ms_per_call": round(1000 * t, 2), "eigvecs_per_s": round(n / t, 1),
                                   "max_residual_n_eps32": float("%.4g" % (res.max() / (n * e32))),
                                   "note": "FP32 host arrays, PCIe staging included (median of 3 calls)"}
    del dS, dP, dZ, V
    torch.cuda.empty_cache()
    return out


def backward_ratio(seed, n=2048, nsel=256, nbatch=32, reps=5):
    """Forward dtgevc3 side='R' howmny='S' (nsel random real eigenvalues, no back-transform) and
    the hand-written backward for nbatch cotangent batches, on the device; times in ms."""
    import torch

    import xtgevc3_amd as X
    from xtgevc3_amd import pencils as G

    S, P, ar, ai, be = G.real_quasi_triangular(n, seed)
    rng = np.random.default_rng(seed)
    real_k = np.array([k for k in range(n) if ai[k] == 0.0 and (k == 0 or ai[k - 1] == 0.0)])
    sel = np.zeros(n, dtype=np.int32)
    sel[np.sort(rng.choice(real_k, size=nsel, replace=False))] = 1
    dS = torch.from_numpy(np.ascontiguousarray(S.T)).cuda()
    dP = torch.from_numpy(np.ascontiguousarray(P.T)).cuda()
    Vt = torch.zeros((nsel, n), dtype=torch.float64, device="cuda")
    Gt = torch.randn((nbatch * nsel, n), dtype=torch.float64, device="cuda")
    gS = torch.zeros((nbatch * n, n), dtype=torch.float64, device="cuda")
    gP = torch.zeros_like(gS)

    def fwd():
        m, info = X.dtgevc3_device("R", "S", sel, dS, dP, ar, ai, be, VR=Vt, mm=nsel)
        assert (m, info) == (nsel, 0), (m, info)

    def bwd():
        X.dtgevc3_backward_device(dS, dP, ar, ai, be, sel, Vt, Gt, nbatch=nbatch, dS=gS, dP=gP)

    G1 = Gt[:nsel].contiguous()
    gS1 = gS[:n].contiguous()
    gP1 = gP[:n].contiguous()

    def bwd1():
        X.dtgevc3_backward_device(dS, dP, ar, ai, be, sel, Vt, G1, nbatch=1, dS=gS1, dP=gP1)

    Vm = torch.zeros_like(Vt)
    gSm = torch.zeros_like(gS)
    gPm = torch.zeros_like(gP)

    def fwd_mixed():
        m, info = X.dtgevc3_device("R", "S", sel, dS, dP, ar, ai, be, VR=Vm, mm=nsel, precision=X.PREC_SPLIT_I8)
        assert (m, info) == (nsel, 0), (m, info)

    def bwd_mixed():
        X.dtgevc3_backward_device(dS, dP, ar, ai, be, sel, Vm, Gt, nbatch=nbatch, dS=gSm, dP=gPm,
                                  precision=X.PREC_SPLIT_I8)

    out = {}
    for name, f in (("fwd", fwd), ("bwd", bwd), ("bwd1", bwd1), ("fwd_mixed", fwd_mixed), ("bwd_mixed", bwd_mixed)):
        f()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            f()
        torch.cuda.synchronize()
        out[name] = 1000.0 * (time.perf_counter() - t) / reps
    # on N(0,1) off-diagonals the exact gradients of most selected eigenvectors exceed the
    # FP64 range at this size (triangular-solve growth): count what is representable
    nonfinite = float((~torch.isfinite(gS)).sum() + (~torch.isfinite(gP)).sum()) / (2.0 * gS.numel())
    return {"n": n, "selected": nsel, "batches": nbatch, "fwd_ms": round(out["fwd"], 3),
            "bwd_ms": round(out["bwd"], 3), "bwd_over_fwd": round(out["bwd"] / out["fwd"], 3),
            "bwd_per_batch_over_fwd": round(out["bwd"] / nbatch / out["fwd"], 4),
            "bwd_one_batch_ms": round(out["bwd1"], 3), "bwd_one_batch_over_fwd": round(out["bwd1"] / out["fwd"], 3),
            "fp64_grad_nonfinite_fraction": round(nonfinite, 4),
            "mixed": {"form": "split-int8", "fwd_ms": round(out["fwd_mixed"], 3),
                      "bwd_ms": round(out["bwd_mixed"], 3),
                      "bwd_over_fwd": round(out["bwd_mixed"] / out["fwd_mixed"], 3)},
            "accuracy": backward_accuracy(seed, n, nsel)}


def backward_accuracy(seed, n=2048, nsel=256, nbatch=4, offdiag=0.003):
    """Gradient error of the mixed forms (forward and backward both mixed) against the FP64
    forward + backward, on a pencil whose FP64 gradients carry digits: the same construction
    with off-diagonals scaled by `offdiag`, the nsel real eigenvalues farthest from all others
    selected.  Errors are max |g - g_fp64| / max |g_fp64| over dL/dS and dL/dP."""
    import torch

    import xtgevc3_amd as X
    from xtgevc3_amd import pencils as G

    S, P, ar, ai, be = G.real_quasi_triangular(n, seed, offdiag=offdiag)
    lam = (ar + 1j * ai) / be
    real_k = np.array([k for k in range(n) if ai[k] == 0.0 and (k == 0 or ai[k - 1] == 0.0)])
    gap = np.array([np.partition(np.abs(lam - lam[k]), 1)[1] for k in real_k])
    pick = np.sort(real_k[np.argsort(-gap)[:nsel]])
    sel = np.zeros(n, dtype=np.int32)
    sel[pick] = 1
    dS = torch.from_numpy(np.ascontiguousarray(S.T)).cuda()
    dP = torch.from_numpy(np.ascontiguousarray(P.T)).cuda()
    gen = torch.Generator("cuda").manual_seed(seed + 1)
    Gt = torch.randn((nbatch * nsel, n), dtype=torch.float64, device="cuda", generator=gen)
    res = {}
    for p in (X.PREC_FP64, X.PREC_SPLIT_I8, X.PREC_MIXED_F32):
        Vt = torch.zeros((nsel, n), dtype=torch.float64, device="cuda")
        m, info = X.dtgevc3_device("R", "S", sel, dS, dP, ar, ai, be, VR=Vt, mm=nsel, precision=p)
        assert (m, info) == (nsel, 0), (m, info)
        gS, gP = X.dtgevc3_backward_device(dS, dP, ar, ai, be, sel, Vt, Gt, nbatch=nbatch, precision=p)
        res[p] = (Vt, torch.cat([gS.flatten(), gP.flatten()]))
    V0, g0 = res[X.PREC_FP64]
    scale = float(g0.abs().max())
    out = {"pencil": f"quasi-triangular pencil (40% pairs), off-diagonals x{offdiag}, {nsel} best-separated real "
                     f"eigenvalues, "
                     f"{nbatch} cotangent batches", "fp64_grad_max": float("%.4g" % scale),
           "fp64_grad_finite": bool(torch.isfinite(g0).all())}
    for p, name in ((X.PREC_SPLIT_I8, "split_int8"), (X.PREC_MIXED_F32, "fp32_products")):
        Vm, gm = res[p]
        out[name] = {"max_rel_grad_err_vs_fp64": float("%.3g" % (float((gm - g0).abs().max()) / scale)),
                     "max_abs_V_err_vs_fp64": float("%.3g" % float((Vm - V0).abs().max()))}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-residual", action="store_true")
    ap.add_argument("--no-backward", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-host-api", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the other configurations (n=512 QZ on CPU and GPU, ztgevc3 n=8192, n=16384 clustered)")
    ap.add_argument("--cluster", action="store_true",
                    help="eigenvalues clustered in 64 groups (relative spread 1e-6), as the n=16384 configuration")
    ap.add_argument("--z-host", action="store_true",
                    help="make Z by a QR on the host (for runs under rocprofv3 counter collection)")
    ap.add_argument("--howmny", default="B", choices=["A", "B"],
                    help="B: back-transformed (the metric); A: eigenvectors of (S, P), no back-transform")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this path runs on one GPU (the eigenvector columns shard across GPUs without "
              f"exchange, but no multi-GPU run is part of this build); got --gpus {args.gpus}", file=sys.stderr)
        sys.exit(2)

    import torch

    import xtgevc3_amd as X

    n = args.n
    dev = "cuda"
    back = args.howmny == "B"
    S, P, ar, ai, be, dS, dP, dZ = make_inputs(n, args.seed, dev, cluster=args.cluster, need_z=back,
                                               z_host=args.z_host)
    dVR = torch.empty((n, n), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()

    def step():
        if back:
            dVR.copy_(dZ)
        m, info = X.dtgevc3_device("R", args.howmny, None, dS, dP, ar, ai, be, VR=dVR, stream=stream.cuda_stream)
        assert info == 0 and m == n, (m, info)
        return m

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log("[bench] warmup done")
    # timed region: K steps, library profiling off
    t0 = time.perf_counter()
    m = 0
    for _ in range(args.steps):
        m += step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    total_s = t1 - t0
    ms_per_step = 1000.0 * total_s / args.steps
    value = m / total_s
    log(f"[bench] {args.steps} steps: {ms_per_step:.2f} ms/step, {value:.1f} eigvecs/s")
    # per-kernel breakdown: separate profiled steps (HIP events around every launch, on the
    # stream the kernels run on), outside the timed region
    nprof = max(1, min(args.steps, 2))
    X.profile_enable(True)
    for _ in range(nprof):
        step()
    torch.cuda.synchronize()
    X.profile_enable(False)
    prof = X.profile_read()
    for k, v in prof.items():
        if v["launches"]:
            avg = v["ms"] / v["launches"]
            log(f"[bench]   {k:14s} launches={v['launches'] // nprof}/step  total={v['ms'] / nprof:.3f} "
                f"ms/step  avg={avg * 1000:.1f} us  {v['flops'] / (v['ms'] * 1e9):.1f} TFLOP/s  "
                f"{v['bytes'] / (v['ms'] * 1e6):.1f} GB/s")

    # algorithmic flops of the whole path: sum_k 2k^2 (solve) + 2n(k+1) (back-transform), pairs on
    # both columns
    kk = np.arange(n, dtype=np.float64)
    flops = float(np.sum(2 * kk * kk) + (np.sum(2 * n * (kk + 1)) if back else 0.0))
    path_tflops = flops * args.steps / total_s / 1e12

    # dominant kernel: the kernel instance with the largest share of the timed region (names as
    # rocprofv3 prints them; at n >= 2048 the back-transform uses the 128-tile instance)
    names = {"solve": "xtg::k_solve_real<65>", "update": "xtg::k_gemm_f64<0, true, 64, 64, 32, false>",
             "update_128x128": "xtg::k_gemm_f64<0, true, 128, 128, 16, false>",
             "update_128x64": "xtg::k_gemm_f64<0, true, 128, 64, 16, false>",
             "update_64x128": "xtg::k_gemm_f64<0, true, 64, 128, 16, false>",
             "back": "xtg::k_gemm_f64<1, true, 128, 128, 16, false>",
             "other": "(copies, scale fix-up, normalisation)"}
    dom = max((k for k in prof if prof[k]["launches"] and k != "other"), key=lambda k: prof[k]["ms"])
    d = prof[dom]
    avg_ms = d["ms"] / d["launches"]
    if dom != "solve":
        # algorithmic flops per launch: 2 M N K of the launch's product (triangular part only
        # for the back-transform)
        ach = d["flops"] / d["launches"] / (avg_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": round(ach, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach / PEAK_FP64_TFLOPS, 4), "traffic": MEASURED_TRAFFIC.get((names[dom], n)),
                "kernel": names[dom],
                "avg_launch_ms": round(avg_ms, 4), "launches_per_step": d["launches"] // nprof}
    else:
        ach = d["bytes"] / d["launches"] / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": MEASURED_TRAFFIC.get((names[dom], n)),
                "kernel": names[dom],
                "avg_launch_ms": round(avg_ms, 4), "launches_per_step": d["launches"] // nprof}
    kernels = {k: {"ms_per_step": round(v["ms"] / nprof, 4), "launches_per_step": v["launches"] // nprof,
                   "tflops": round(v["flops"] / (v["ms"] * 1e9), 3) if v["ms"] else None}
               for k, v in prof.items() if v["launches"]}

    res = None
    if not args.no_residual:
        step()
        torch.cuda.synchronize()
        res, ncols = residual_every_vector(dS, dP, ar, ai, be, dVR, args.howmny, dZ)
        log(f"[bench] residual (metric A) of every one of the {ncols} vectors: max {res:.3e} n*eps")

    # mixed forms on the same inputs: split-int8 digit products (FP64-level residual) and FP32
    # products; each timed over the same K steps, then one profiled step and one residual step
    def run_form(precision, label):
        def step_form():
            if back:
                dVR.copy_(dZ)
            m, info = X.dtgevc3_device("R", args.howmny, None, dS, dP, ar, ai, be, VR=dVR, stream=stream.cuda_stream,
                                       precision=precision)
            assert info == 0 and m == n, (m, info)
            return m

        step_form()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        mm_ = 0
        for _ in range(args.steps):
            mm_ += step_form()
        torch.cuda.synchronize()
        tm = time.perf_counter() - t0
        out = {"form": label, "value": round(mm_ / tm, 2), "ms_per_step": round(1000.0 * tm / args.steps, 3),
               "speedup_vs_plain": round(total_s / tm, 3)}
        X.profile_enable(True)
        step_form()
        torch.cuda.synchronize()
        X.profile_enable(False)
        pf = X.profile_read()
        out["kernels"] = {k: {"ms_per_step": round(v["ms"], 4), "launches_per_step": v["launches"],
                              "tflops": round(v["flops"] / (v["ms"] * 1e9), 3) if v["ms"] else None}
                          for k, v in pf.items() if v["launches"]}
        for k, v in out["kernels"].items():
            log(f"[bench]   {label[:10]:10s} {k:14s} launches={v['launches_per_step']}  {v['ms_per_step']:.3f} ms  "
                f"{v['tflops']} TFLOP/s")
        if "back_split" in pf and pf["back_split"]["launches"]:
            # the split form's back-transform kernel against the int8 MFMA peak: each FP64-equivalent
            # multiply-add costs SPLIT_PRODUCTS int8 digit products
            v = pf["back_split"]
            ach = v["flops"] * SPLIT_PRODUCTS / (v["ms"] * 1e9)
            out["roofline"] = {"bound": "mfma", "achieved": round(ach, 1), "peak": PEAK_INT8_TOPS, "unit": "TOP/s",
                               "frac": round(ach / PEAK_INT8_TOPS, 4),
                               "traffic": MEASURED_TRAFFIC.get(("xtg::k_gemm_i8<128, 128, 2, 4> (back-transform)", n)),
                               "kernel": "xtg::k_gemm_i8<128, 128, 2, 4> (back-transform)",
                               "avg_launch_ms": round(v["ms"] / v["launches"], 4),
                               "fp64_equivalent_tflops": round(v["flops"] / (v["ms"] * 1e9), 2)}
        if not args.no_residual:
            step()
            torch.cuda.synchronize()
            ref64 = dVR.clone()
            step_form()
            torch.cuda.synchronize()
            rm, _ = residual_every_vector(dS, dP, ar, ai, be, dVR, args.howmny, dZ)
            diff = float((dVR - ref64).abs().max())
            out["max_residual_n_eps"] = float(f"{rm:.4g}")
            out["max_abs_diff_vs_fp64"] = float(f"{diff:.4g}")
            del ref64
        log(f"[bench] {label}: {out['ms_per_step']:.2f} ms/step, {out['value']:.1f} eigvecs/s, "
            f"x{out['speedup_vs_plain']} vs plain; residual {out.get('max_residual_n_eps')} n*eps, "
            f"max |V - V_fp64| {out.get('max_abs_diff_vs_fp64')}")
        return out

    mixed = mixed_f32 = None
    if not args.no_mixed:
        mixed = run_form(X.PREC_SPLIT_I8, "split-int8 digit products (5 base-255 digits, exact int32 "
                                          "accumulation; far updates, back-transform), FP64 solves/near "
                                          "updates/scales/normalisation")
        mixed_f32 = run_form(X.PREC_MIXED_F32, "FP32 MFMA products (updates, back-transform), FP64 "
                                               "solves/scales/normalisation")

    # the drop-in host-pointer interface (reference signature): S, P, Z staged over PCIe, VR
    # copied back; PCIe-inclusive rate (median of 3 calls: host-side packing of pageable memory
    # varies from call to call; not the reported value), also for the FP32 interface stgevc3 on
    # the same pencil
    host = None
    if not args.no_host_api and back:
        Sf = np.asfortranarray(S)
        Pf = np.asfortranarray(P)
        Zf = np.asfortranarray(dZ.t().cpu().numpy())
        host = {}
        for name, fn, dt in (("dtgevc3", X.dtgevc3, np.float64), ("stgevc3", X.stgevc3, np.float32)):
            Sx, Px = np.asfortranarray(Sf.astype(dt)), np.asfortranarray(Pf.astype(dt))
            Vx = np.asfortranarray(Zf.astype(dt))
            fn("R", "B", None, Sx, Px, ar.astype(dt), ai.astype(dt), be.astype(dt), VR=Vx.copy(order="F"))
            times = []
            for _ in range(3):
                Vx2 = Vx.copy(order="F")
                t0 = time.perf_counter()
                mh, ih, _ = fn("R", "B", None, Sx, Px, ar.astype(dt), ai.astype(dt), be.astype(dt), VR=Vx2)
                times.append(time.perf_counter() - t0)
                assert ih == 0 and mh == n, (mh, ih)
                del Vx2
            th = sorted(times)[1]
            host[name] = {"ms_per_call": round(1000.0 * th, 2), "eigvecs_per_s": round(mh / th, 1),
                          "ms_calls": [round(1000.0 * t, 2) for t in times]}
            log(f"[bench] host-pointer {name} (PCIe staging included): {1000.0 * th:.1f} ms median of "
                f"{', '.join(f'{1000.0 * t:.1f}' for t in times)}, {mh / th:.1f} eigvecs/s")
        del Sf, Pf, Zf

    bw = None
    if not args.no_backward:
        bw = backward_ratio(args.seed)
        log(f"[bench] backward n={bw['n']} sel={bw['selected']} x{bw['batches']} batches: fwd {bw['fwd_ms']} ms, "
            f"bwd {bw['bwd_ms']} ms (all batches), ratio {bw['bwd_over_fwd']} ({bw['bwd_per_batch_over_fwd']} per batch); "
            f"one batch {bw['bwd_one_batch_ms']} ms, ratio {bw['bwd_one_batch_over_fwd']}; "
            f"split-int8 fwd {bw['mixed']['fwd_ms']} ms, bwd {bw['mixed']['bwd_ms']} ms; "
            f"FP64 gradient non-finite fraction {bw['fp64_grad_nonfinite_fraction']}; accuracy {bw['accuracy']}")

    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        cpu = cpu_baseline(n, args.seed)
        if cpu:
            log(f"[bench] cpu baseline {cpu['value']:.1f} eigvecs/s ({cpu['sample']})")
            legs = cpu_legs(n, args.seed)
            if legs:
                legs["cfg0_dtgevc3_B_B_n512_qz"] = config0_cpu(args.seed)
                cpu["legs"] = legs
                log(f"[bench] cpu legs {legs}")

    extra = None
    if not args.no_extra:
        extra = {}
        for key, fn in (("cfg0_dtgevc3_B_B_n512_qz", lambda: config0(args.seed)),
                        ("cfg2_ztgevc3_B_B_n8192", lambda: config2(args.seed, max(2, min(args.steps, 3)))),
                        ("cfg3_dtgevc3_R_n16384_clustered", lambda: config3(args.seed, max(2, min(args.steps, 3))))):
            t0 = time.perf_counter()
            extra[key] = fn()
            log(f"[bench] {key} ({time.perf_counter() - t0:.1f} s): {json.dumps(extra[key])}")

    out = {
        "metric": "eigvecs/s at n=8192: mixed vs plain FP64 GPU at fixed residual; bwd/fwd ratio",
        "value": round(value, 2),
        "unit": "eigvecs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"dtgevc3 side=R howmny={args.howmny} n={n} plain FP64, quasi-triangular pencil "
                               f"(40% 2x2 pairs)" + (", 64 eigenvalue clusters" if args.cluster else "")
                               + (", Z orthogonal" if back else ""), "n": n, "side": "R",
                   "howmny": args.howmny,
                   "form": "plain-fp64", "path_tflops": round(path_tflops, 3),
                   "max_residual_n_eps": None if res is None else float(f"{res:.4g}"),
                   "residual_columns": "every vector"},
        "roofline": roof,
        "mixed": mixed,
        "mixed_f32": mixed_f32,
        "host_api": host,
        "backward": bw,
        "kernels": kernels,
        "cpu_baseline": None if cpu is None else {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample",
                                                                      "legs") if k in cpu},
        "other_configs": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
