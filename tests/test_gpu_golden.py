"""The committed golden fixtures (tests/golden/*.npz, written by the reference library compiled
from its own sources) through the HIP path, host-pointer and device-pointer entries.

They hold the reference tests' own edge cases: the 4x4 known-answer pencils
(tests/test_dtgevc3.cpp:318-434, tests/test_stgevc3.cpp:236-313, tests/test_ztgevc3.cpp:146-201),
the infinite eigenvalue (beta = 0; beta subnormal in the FP32 / complex64 interfaces, which
stgevc3 / ctgevc3 count as infinite below FLT_MIN), the 1e+-150 and 1e+-30 scalings, and the
six bignum / smlnum scalings of tests/test_lapack_d.cpp:328-344 and tests/test_lapack_z.cpp:248-264.
m and info must match bit for bit; the vectors, normalised the LAPACK way on both sides,
within 100 * eps * cond_k (cond_k as in tests/helpers.eigvec_cond_sampled)."""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import EPS, compare_sampled, eigvec_cond_sampled

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _names(*prefixes):
    out = []
    for p in prefixes:
        out += sorted(os.path.basename(f)[:-4] for f in glob.glob(os.path.join(GOLD, p + "*.npz")))
    return out


def _load(name):
    return dict(np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False))


def _sel(g):
    return g["sel"].astype(np.int32) if "sel" in g and g["sel"].size else None


def _real_expected(g, m, kind):
    return [O.normalize(np.asfortranarray(g[k][:, :m].astype(np.float64)), kind) for k in ("VL", "VR")]


def _check_real(g, gL, gR, m, kind, idx, tol_factor=100.0, eps=EPS):
    eL, eR = _real_expected(g, m, kind)
    al, bb = O.eig_alpha_beta(g["ar"], g["ai"], g["be"], idx)
    for name, mine, ref, left in (("L", gL, eL, True), ("R", gR, eR, False)):
        mine = np.asarray(mine[:, :m], dtype=np.float64)
        assert np.all(np.isfinite(mine)), name
        cond = eigvec_cond_sampled(g["S"], g["P"], al, bb, kind, idx, left=left)
        bad, err, tol = compare_sampled(mine, ref, cond, tol_factor * eps / EPS)
        assert bad.size == 0, f"{name}: cols {bad[:8]} err {err[bad[:8]]} tol {tol[bad[:8]]}"


@pytest.mark.parametrize("name", _names("kat_d", "rand_d", "scale_d"))
def test_golden_dtgevc3_host(name):
    import xtgevc3_amd as X

    g = _load(name)
    n = g["S"].shape[0]
    howmny, side = str(g["howmny"]), str(g["side"])
    sel = _sel(g)
    vin = g.get("Vin", np.eye(n)) if howmny == "B" else np.zeros((n, n))
    VL = np.asfortranarray(vin.astype(np.float64).copy())
    VR = np.asfortranarray(vin.astype(np.float64).copy())
    m, info, _ = X.dtgevc3(side, howmny, sel, np.asfortranarray(g["S"]), np.asfortranarray(g["P"]), g["ar"], g["ai"],
                           g["be"], VL=VL, VR=VR)
    assert (m, info) == (int(g["m"]), int(g["info"]))
    kind, idx = O.real_layout(howmny, sel, g["ai"])
    _check_real(g, VL, VR, m, kind, idx)


@pytest.mark.parametrize("name", _names("kat_d", "scale_d"))
@pytest.mark.parametrize("precision", ["fp64", "split_i8"])
def test_golden_dtgevc3_device(name, precision):
    """The same fixtures through xtgevc3_dtgevc3_device(_ex); the split-int8 form is held to
    its documented floor of max(10 n eps, 2.5e-11) relative error in place of 100 eps."""
    import torch

    import xtgevc3_amd as X

    g = _load(name)
    n = g["S"].shape[0]
    prec = X.PREC_FP64 if precision == "fp64" else X.PREC_SPLIT_I8
    dS, dP = X.fortran_device(g["S"]), X.fortran_device(g["P"])
    dVL, dVR = X.fortran_device(np.eye(n)), X.fortran_device(np.eye(n))
    m, info = X.dtgevc3_device("B", "B", None, dS, dP, g["ar"], g["ai"], g["be"], VL=dVL, VR=dVR, precision=prec)
    torch.cuda.synchronize()
    assert (m, info) == (int(g["m"]), int(g["info"]))
    kind, idx = O.real_layout("B", None, g["ai"])
    eps = EPS if precision == "fp64" else max(10 * n * EPS, 2.5e-11) / 100.0
    _check_real(g, X.from_fortran_device(dVL), X.from_fortran_device(dVR), m, kind, idx, eps=eps)


@pytest.mark.parametrize("name", ["kat_s_big", "kat_s_small", "kat_s_inf", "kat_s_sub"])
def test_golden_stgevc3(name):
    """FP32 interface: computed in FP64 on the device and rounded, so within FP32 rounding of
    the reference's FP32 result; the subnormal beta is an infinite eigenvalue as in stgevc3."""
    import xtgevc3_amd as X

    g = _load(name)
    VL = np.asfortranarray(np.eye(4, dtype=np.float32))
    VR = np.asfortranarray(np.eye(4, dtype=np.float32))
    m, info, _ = X.stgevc3("B", "B", None, np.asfortranarray(g["S"]), np.asfortranarray(g["P"]), g["ar"], g["ai"],
                           g["be"], VL=VL, VR=VR)
    assert (m, info) == (int(g["m"]), int(g["info"]))
    kind, idx = O.real_layout("B", None, g["ai"].astype(np.float64))
    for mine, key in ((VL, "VL"), (VR, "VR")):
        ref = O.normalize(np.asfortranarray(g[key].astype(np.float32)), kind)
        assert np.all(np.isfinite(mine))
        assert np.max(np.abs(mine - ref)) < 64 * np.finfo(np.float32).eps, key


def _check_cplx(g, gL, gR, m, dt=np.complex128, tol=None):
    n = g["S"].shape[0]
    sel = _sel(g)
    howmny = str(g["howmny"])
    idx = np.arange(n) if sel is None or howmny != "S" else np.nonzero(sel)[0]
    kind = np.zeros(m, dtype=np.int32)
    al, bb = g["al"].astype(np.complex128)[idx], g["be"].astype(np.complex128)[idx]
    for name, mine, key, left in (("L", gL, "VL", True), ("R", gR, "VR", False)):
        ref = O.normalize(np.asfortranarray(g[key][:, :m].astype(dt)))
        mine = mine[:, :m]
        assert np.all(np.isfinite(mine)), name
        if tol is not None:
            assert np.max(np.abs(mine - ref)) < tol, name
            continue
        cond = eigvec_cond_sampled(g["S"], g["P"], al, bb, kind, idx, left=left)
        bad, err, tl = compare_sampled(mine, ref, cond)
        assert bad.size == 0, f"{name}: cols {bad[:8]} err {err[bad[:8]]} tol {tl[bad[:8]]}"


@pytest.mark.parametrize("name", _names("kat_z", "rand_z", "scale_z"))
def test_golden_ztgevc3_host(name):
    import xtgevc3_amd as X

    g = _load(name)
    n = g["S"].shape[0]
    howmny, side = str(g["howmny"]), str(g["side"])
    vin = g.get("Vin", np.eye(n)) if howmny == "B" else np.zeros((n, n))
    VL = np.asfortranarray(vin.astype(np.complex128).copy())
    VR = np.asfortranarray(vin.astype(np.complex128).copy())
    m, info, _ = X.ztgevc3(side, howmny, _sel(g), np.asfortranarray(g["S"]), np.asfortranarray(g["P"]), g["al"],
                           g["be"], VL=VL, VR=VR)
    assert (m, info) == (int(g["m"]), int(g["info"]))
    _check_cplx(g, VL, VR, m)


@pytest.mark.parametrize("name", _names("kat_z", "scale_z"))
def test_golden_ztgevc3_device(name):
    import torch

    import xtgevc3_amd as X

    g = _load(name)
    n = g["S"].shape[0]
    eye = np.eye(n, dtype=np.complex128)
    dS, dP = X.fortran_device(g["S"]), X.fortran_device(g["P"])
    dVL, dVR = X.fortran_device(eye), X.fortran_device(eye)
    m, info = X.ztgevc3_device("B", "B", None, dS, dP, g["al"], g["be"], VL=dVL, VR=dVR)
    torch.cuda.synchronize()
    assert (m, info) == (int(g["m"]), int(g["info"]))
    _check_cplx(g, X.from_fortran_device(dVL), X.from_fortran_device(dVR), m)


def test_golden_ctgevc3_subnormal_beta():
    """complex64 interface, beta_3 = 1e-40 (subnormal in FP32): an infinite eigenvalue as in
    ctgevc3 (|beta| <= FLT_MIN)."""
    import xtgevc3_amd as X

    g = _load("kat_c_sub")
    VL = np.asfortranarray(np.eye(4, dtype=np.complex64))
    VR = np.asfortranarray(np.eye(4, dtype=np.complex64))
    m, info, _ = X.ctgevc3("B", "B", None, np.asfortranarray(g["S"]), np.asfortranarray(g["P"]), g["al"], g["be"],
                           VL=VL, VR=VR)
    assert (m, info) == (int(g["m"]), int(g["info"]))
    _check_cplx(g, VL, VR, m, dt=np.complex64, tol=64 * np.finfo(np.float32).eps)
