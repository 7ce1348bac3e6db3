"""xtgevc3_amd -- MI355X (gfx950) implementation of the blocked xTGEVC3 eigenvector path.

Python mirror of the reference's operator interface (vlovero/xTGEVC3 include/tgevc3.h:10-13):
``dtgevc3``/``stgevc3``/``ztgevc3``/``ctgevc3`` take the same arguments with the same
meaning (side, howmny, select, S, P, eigenvalues, VL, VR, mm, lwork) and return
``(m, info, work0)`` where ``work0`` is ``work[0]`` after the call (the workspace size on a
``lwork=-1`` query).  VL/VR are Fortran-ordered numpy arrays updated in place, as the
reference updates its VL/VR arrays.

``*_device`` variants take torch tensors already resident on the GPU (column-major
matrices passed as their transposed row-major storage, see ``_fortran``).

Everything calls into ``libxtgevc3.so`` through its C ABI (include/xtgevc3.h).  There is no
CPU fallback: if the library is missing, import fails.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libxtgevc3.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
                      "or `make -C xtgevc3_amd/csrc`")

# torch ships its own copy of the HIP runtime (same soname, libamdhip64.so.7).  Loading torch
# first lets this library bind to that copy, so a process that uses both talks to the GPU
# through one runtime.
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover - torch is only plumbing
    torch = None

_lib = ctypes.CDLL(LIB_PATH)
_lib.xtgevc3_last_error.restype = ctypes.c_char_p
_lib.xtgevc3_abi_version.restype = ctypes.c_int

INFO_DEVICE_ERROR = -1000
INFO_UNSUPPORTED = -1001
PREC_FP64, PREC_MIXED_F32, PREC_SPLIT_I8 = 0, 1, 3

EXPORTED_SYMBOLS = [
    "xtgevc3_abi_version", "xtgevc3_last_error",
    "xtgevc3_dtgevc3", "xtgevc3_stgevc3", "xtgevc3_ztgevc3", "xtgevc3_ctgevc3",
    "xtgevc3_dtgevc3_device", "xtgevc3_dtgevc3_device_ex", "xtgevc3_ztgevc3_device", "xtgevc3_ztgevc3_device_ex",
    "xtgevc3_profile_enable", "xtgevc3_profile_read", "xtgevc3_dtgevc3_backward_device",
    "xtgevc3_dtgevc3_backward_device_ex", "xtgevc3_dtgsna_e_device", "xtgevc3_ztgsna_e_device",
    "xtgevc3_dtgevc3_device_noalloc", "xtgevc3_dtgevc3_backward_full_device",
]

NOALLOC_MAX_N = 512

PROF_KINDS = ("solve", "update", "back", "other", "update_128x128", "update_128x64", "update_64x128",
              "update_split", "back_split", "split_digits")


def profile_enable(on=True):
    _lib.xtgevc3_profile_enable(ctypes.c_int(1 if on else 0))


def profile_read():
    """-> {kind: dict(ms=total, launches=, flops=, bytes=)}; synchronises and resets."""
    k = len(PROF_KINDS)
    ms = (ctypes.c_double * k)()
    la = (ctypes.c_long * k)()
    fl = (ctypes.c_double * k)()
    by = (ctypes.c_double * k)()
    _lib.xtgevc3_profile_read(ms, la, fl, by)
    return {PROF_KINDS[i]: dict(ms=ms[i], launches=la[i], flops=fl[i], bytes=by[i]) for i in range(k)}


def lib():
    return _lib


def last_error():
    return _lib.xtgevc3_last_error().decode()


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


_REAL = {"d": np.float64, "s": np.float32}
_CPLX = {"z": np.complex128, "c": np.complex64}


def _ld(a):
    return max(1, a.shape[0]) if a is not None else 1


def tgevc3(prefix, side, howmny, select, S, P, alpha, beta, alphai=None, VL=None, VR=None, mm=None, lwork=None,
           work=None):
    """Call ``xtgevc3_{prefix}tgevc3`` (host pointers) with the reference's argument meaning.

    For the real prefixes ('d', 's') ``alpha`` is alphar and ``alphai`` must be given.
    ``lwork=None`` performs the reference's two-call idiom (query, then call).  Returns
    (m, info, work0)."""
    real = prefix in _REAL
    dt = _REAL[prefix] if real else _CPLX[prefix]
    fn = getattr(_lib, f"xtgevc3_{prefix}tgevc3")
    n = S.shape[0]
    assert S.dtype == dt and P.dtype == dt and S.flags.f_contiguous and P.flags.f_contiguous
    for V in (VL, VR):
        if V is not None:
            assert V.dtype == dt and V.flags.f_contiguous
    if mm is None:
        mm = n
    sel = None if select is None else np.ascontiguousarray(select, dtype=np.int32)
    m = ctypes.c_int(-1)
    info = ctypes.c_int(-99)
    if lwork is None:
        q = np.zeros(1, dtype=dt)
        _invoke(fn, real, side, howmny, sel, n, S, P, alpha, alphai, beta, dt, VL, VR, mm, m, q, -1, info)
        if info.value != 0 or n == 0:
            return m.value, info.value, q[0]
        lwork = int(np.real(q[0]))
    if work is None:
        work = np.zeros(max(lwork, 1), dtype=dt)
    _invoke(fn, real, side, howmny, sel, n, S, P, alpha, alphai, beta, dt, VL, VR, mm, m, work, lwork, info)
    if info.value == INFO_DEVICE_ERROR:
        raise RuntimeError("xtgevc3 device path failed: " + last_error())
    return m.value, info.value, work[0]


def _invoke(fn, real, side, howmny, sel, n, S, P, alpha, alphai, beta, dt, VL, VR, mm, m, work, lwork, info):
    c = ctypes
    if real:
        ar = np.ascontiguousarray(alpha, dtype=dt)
        ai = np.ascontiguousarray(alphai, dtype=dt)
        be = np.ascontiguousarray(beta, dtype=dt)
        ev = (_ptr(ar), _ptr(ai), _ptr(be))
    else:
        al = np.ascontiguousarray(alpha, dtype=dt)
        be = np.ascontiguousarray(beta, dtype=dt)
        ev = (_ptr(al), _ptr(be))
    fn(c.c_char(side.encode()), c.c_char(howmny.encode()), _ptr(sel), c.c_int(n), _ptr(S), c.c_int(_ld(S)),
       _ptr(P), c.c_int(_ld(P)), *ev, _ptr(VL), c.c_int(_ld(VL)), _ptr(VR), c.c_int(_ld(VR)), c.c_int(mm),
       c.byref(m), _ptr(work), c.c_int(lwork), c.byref(info))


def dtgevc3(side, howmny, select, S, P, alphar, alphai, beta, VL=None, VR=None, mm=None, lwork=None, work=None):
    return tgevc3("d", side, howmny, select, S, P, alphar, beta, alphai=alphai, VL=VL, VR=VR, mm=mm, lwork=lwork,
                  work=work)


def stgevc3(side, howmny, select, S, P, alphar, alphai, beta, VL=None, VR=None, mm=None, lwork=None, work=None):
    return tgevc3("s", side, howmny, select, S, P, alphar, beta, alphai=alphai, VL=VL, VR=VR, mm=mm, lwork=lwork,
                  work=work)


def ztgevc3(side, howmny, select, S, P, alpha, beta, VL=None, VR=None, mm=None, lwork=None, work=None):
    return tgevc3("z", side, howmny, select, S, P, alpha, beta, VL=VL, VR=VR, mm=mm, lwork=lwork, work=work)


def ctgevc3(side, howmny, select, S, P, alpha, beta, VL=None, VR=None, mm=None, lwork=None, work=None):
    return tgevc3("c", side, howmny, select, S, P, alpha, beta, VL=VL, VR=VR, mm=mm, lwork=lwork, work=work)


# ---------------------------------------------------------------------------------------
# device-pointer variants (torch tensors on the GPU)
# ---------------------------------------------------------------------------------------
def _tptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def fortran_device(a_np, device="cuda"):
    """Column-major numpy matrix -> torch tensor on the device holding the same column-major
    storage (returned as the (cols, rows) row-major tensor == its transpose)."""
    import torch

    return torch.from_numpy(np.ascontiguousarray(np.asarray(a_np).T)).to(device)


def from_fortran_device(t):
    return np.asfortranarray(t.cpu().numpy().T)


def tgevc3_device(prefix, side, howmny, select, S, P, alpha, beta, alphai=None, VL=None, VR=None, mm=None,
                  stream=None, precision=PREC_FP64):
    """Device-pointer call.  S, P, VL, VR are torch tensors holding column-major storage
    (row-major tensors of shape (cols, ld)); select and eigenvalues are host numpy arrays.
    precision (real interface only): PREC_FP64, PREC_MIXED_F32 or PREC_SPLIT_I8.  Returns (m, info)."""
    import torch

    c = ctypes
    real = prefix in _REAL
    dt = _REAL[prefix] if real else _CPLX[prefix]
    if precision != PREC_FP64 and prefix not in ("d", "z"):
        raise ValueError("the mixed forms are available for dtgevc3 and ztgevc3 only")
    fn = getattr(_lib, f"xtgevc3_{prefix}tgevc3_device" + ("_ex" if precision != PREC_FP64 else ""))
    n = S.shape[0]
    lds = S.shape[1]
    ldp = P.shape[1]
    ldvl = VL.shape[1] if VL is not None else 1
    ldvr = VR.shape[1] if VR is not None else 1
    if mm is None:
        mm = n
    sel = None if select is None else np.ascontiguousarray(select, dtype=np.int32)
    m = c.c_int(-1)
    info = c.c_int(-99)
    if real:
        ev = [np.ascontiguousarray(alpha, dtype=dt), np.ascontiguousarray(alphai, dtype=dt),
              np.ascontiguousarray(beta, dtype=dt)]
    else:
        ev = [np.ascontiguousarray(alpha, dtype=dt), np.ascontiguousarray(beta, dtype=dt)]
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    args = [c.c_char(side.encode()), c.c_char(howmny.encode()), _ptr(sel), c.c_int(n), _tptr(S), c.c_int(lds),
            _tptr(P), c.c_int(ldp), *[_ptr(e) for e in ev], _tptr(VL), c.c_int(ldvl), _tptr(VR), c.c_int(ldvr),
            c.c_int(mm), c.byref(m), c.byref(info), c.c_void_p(stream)]
    if precision != PREC_FP64:
        args.append(c.c_int(precision))
    rc = fn(*args)
    if rc == INFO_UNSUPPORTED:
        raise ValueError("xtgevc3: " + last_error())
    if rc != 0:
        raise RuntimeError("xtgevc3 device path failed: " + last_error())
    return m.value, info.value


def dtgevc3_device_noalloc(side, howmny, select, S, P, alphar, alphai, beta, VL=None, VR=None, mm=None,
                           stream=None):
    """``xtgevc3_dtgevc3_device_noalloc``: the FP64 device call with no device allocation (a
    static per-GPU workspace, n <= NOALLOC_MAX_N); arguments and result as dtgevc3_device.
    Raises ValueError past the size limit."""
    import torch

    c = ctypes
    n = S.shape[0]
    if mm is None:
        mm = n
    sel = None if select is None else np.ascontiguousarray(select, dtype=np.int32)
    ev = [np.ascontiguousarray(x, dtype=np.float64) for x in (alphar, alphai, beta)]
    m = c.c_int(-1)
    info = c.c_int(-99)
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    rc = _lib.xtgevc3_dtgevc3_device_noalloc(
        c.c_char(side.encode()), c.c_char(howmny.encode()), _ptr(sel), c.c_int(n), _tptr(S), c.c_int(S.shape[1]),
        _tptr(P), c.c_int(P.shape[1]), *[_ptr(e) for e in ev], _tptr(VL), c.c_int(VL.shape[1] if VL is not None else 1),
        _tptr(VR), c.c_int(VR.shape[1] if VR is not None else 1), c.c_int(mm), c.byref(m), c.byref(info),
        c.c_void_p(stream))
    if rc == INFO_UNSUPPORTED:
        raise ValueError("xtgevc3: " + last_error())
    if rc != 0:
        raise RuntimeError("xtgevc3 device path failed: " + last_error())
    return m.value, info.value


def dtgevc3_device(side, howmny, select, S, P, alphar, alphai, beta, VL=None, VR=None, mm=None, stream=None,
                   precision=PREC_FP64):
    return tgevc3_device("d", side, howmny, select, S, P, alphar, beta, alphai=alphai, VL=VL, VR=VR, mm=mm,
                         stream=stream, precision=precision)


def ztgevc3_device(side, howmny, select, S, P, alpha, beta, VL=None, VR=None, mm=None, stream=None,
                   precision=PREC_FP64):
    return tgevc3_device("z", side, howmny, select, S, P, alpha, beta, VL=VL, VR=VR, mm=mm, stream=stream,
                         precision=precision)


def dtgsna_e_device(howmny, select, S, P, alphai, VL, VR, m=None, s=None, stream=None):
    """Eigenvalue condition numbers (LAPACK dtgsna, JOB='E') of the eigenvalues howmny / select
    describes, from their left and right eigenvectors VL, VR as dtgevc3 returns them.  S, P, VL,
    VR: torch tensors holding column-major storage (shape (cols, ld)); select, alphai: host
    arrays.  Returns a float64 device tensor of m values (one per column)."""
    import torch

    c = ctypes
    n = S.shape[0]
    if m is None:
        m = VR.shape[0]
    if s is None:
        s = torch.empty(max(m, 1), dtype=torch.float64, device=S.device)
    sel = None if select is None else np.ascontiguousarray(select, dtype=np.int32)
    ai = np.ascontiguousarray(alphai, dtype=np.float64)
    info = c.c_int(-99)
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    rc = _lib.xtgevc3_dtgsna_e_device(
        c.c_char(howmny.encode()), _ptr(sel), c.c_int(n), _tptr(S), c.c_int(S.shape[1]), _tptr(P), c.c_int(P.shape[1]),
        _ptr(ai), _tptr(VL), c.c_int(VL.shape[1]), _tptr(VR), c.c_int(VR.shape[1]), c.c_int(m), _tptr(s),
        c.byref(info), c.c_void_p(stream))
    if rc != 0:
        raise RuntimeError("xtgevc3 dtgsna_e failed: " + last_error())
    return s[:m]


def ztgsna_e_device(howmny, select, S, P, VL, VR, m=None, s=None, stream=None):
    """Eigenvalue condition numbers of a complex pencil (LAPACK ztgsna, JOB='E'); arguments as
    dtgsna_e_device with complex128 tensors and no alphai."""
    import torch

    c = ctypes
    n = S.shape[0]
    if m is None:
        m = VR.shape[0]
    if s is None:
        s = torch.empty(max(m, 1), dtype=torch.float64, device=S.device)
    sel = None if select is None else np.ascontiguousarray(select, dtype=np.int32)
    info = c.c_int(-99)
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    rc = _lib.xtgevc3_ztgsna_e_device(
        c.c_char(howmny.encode()), _ptr(sel), c.c_int(n), _tptr(S), c.c_int(S.shape[1]), _tptr(P), c.c_int(P.shape[1]),
        _tptr(VL), c.c_int(VL.shape[1]), _tptr(VR), c.c_int(VR.shape[1]), c.c_int(m), _tptr(s), c.byref(info),
        c.c_void_p(stream))
    if rc != 0:
        raise RuntimeError("xtgevc3 ztgsna_e failed: " + last_error())
    return s[:m]


def dtgevc3_backward_device(S, P, alphar, alphai, beta, select, V, G, nbatch=1, dS=None, dP=None, stream=None,
                            precision=PREC_FP64, return_info=False):
    """Backward of ``dtgevc3_device('R', 'S' or 'A', ...)`` for real selected eigenvalues.
    S, P, V, G, dS, dP: torch tensors holding column-major storage (shape (cols, ld)).
    G holds nbatch * m columns.  Returns (dS, dP) with nbatch stacked n x n matrices each
    (tensors of shape (nbatch * n, n) in column-major storage).  precision: PREC_FP64,
    PREC_SPLIT_I8 or PREC_MIXED_F32 (far updates of the adjoint sweep in that form).  A positive info
    (adjoints past the double range, their gradient entries NaN) gives a RuntimeWarning, and
    with return_info=True the result is (dS, dP, info)."""
    import torch

    c = ctypes
    n = S.shape[0]
    if dS is None:
        dS = torch.zeros((nbatch * n, n), dtype=torch.float64, device=S.device)
    if dP is None:
        dP = torch.zeros((nbatch * n, n), dtype=torch.float64, device=S.device)
    sel = None if select is None else np.ascontiguousarray(select, dtype=np.int32)
    ev = [np.ascontiguousarray(x, dtype=np.float64) for x in (alphar, alphai, beta)]
    info = c.c_int(-99)
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    rc = _lib.xtgevc3_dtgevc3_backward_device_ex(
        c.c_int(n), _tptr(S), c.c_int(S.shape[1]), _tptr(P), c.c_int(P.shape[1]), *[_ptr(e) for e in ev], _ptr(sel),
        _tptr(V), c.c_int(V.shape[1]), _tptr(G), c.c_int(G.shape[1]), c.c_int(nbatch), _tptr(dS), c.c_int(n),
        _tptr(dP), c.c_int(n), c.byref(info), c.c_void_p(stream), c.c_int(precision))
    if rc == INFO_UNSUPPORTED:
        raise ValueError("xtgevc3: " + last_error())
    if rc != 0:
        raise RuntimeError("xtgevc3 backward failed: " + last_error())
    if info.value < 0:
        raise ValueError(f"xtgevc3 backward: info={info.value} (complex or infinite selected eigenvalue)")
    if info.value > 0:
        import warnings

        warnings.warn(f"xtgevc3 backward: {info.value} adjoint vectors exceed the double range; the gradient "
                      "entries that depend on them are NaN", RuntimeWarning, stacklevel=2)
    if return_info:
        return dS, dP, info.value
    return dS, dP


def dtgevc3_backward_full_device(side, S, P, alphar, alphai, beta, select, V, G, nbatch=1, Z=None, dZ=None,
                                 stream=None, return_info=False):
    """Backward for side 'R' or 'L' (left eigenvectors), optionally through the back-transform:
    with Z (Z for right, Q for left vectors) V holds the back-transformed vectors and dZ
    (allocated here when not given) receives dL/dZ.  Tensors hold column-major storage as in
    dtgevc3_backward_device.  Returns (dS, dP) or (dS, dP, dZ) when Z is given (and info last
    with return_info=True)."""
    import torch

    c = ctypes
    n = S.shape[0]
    dS = torch.zeros((nbatch * n, n), dtype=torch.float64, device=S.device)
    dP = torch.zeros((nbatch * n, n), dtype=torch.float64, device=S.device)
    if Z is not None and dZ is None:
        dZ = torch.zeros((nbatch * n, n), dtype=torch.float64, device=S.device)
    sel = None if select is None else np.ascontiguousarray(select, dtype=np.int32)
    ev = [np.ascontiguousarray(x, dtype=np.float64) for x in (alphar, alphai, beta)]
    info = c.c_int(-99)
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    rc = _lib.xtgevc3_dtgevc3_backward_full_device(
        c.c_char(side.encode()), c.c_int(n), _tptr(S), c.c_int(S.shape[1]), _tptr(P), c.c_int(P.shape[1]),
        *[_ptr(e) for e in ev], _ptr(sel), _tptr(Z), c.c_int(Z.shape[1] if Z is not None else 1), _tptr(V),
        c.c_int(V.shape[1]), _tptr(G), c.c_int(G.shape[1]), c.c_int(nbatch), _tptr(dS), c.c_int(n), _tptr(dP),
        c.c_int(n), _tptr(dZ), c.c_int(n), c.byref(info), c.c_void_p(stream))
    if rc != 0:
        raise RuntimeError("xtgevc3 backward failed: " + last_error())
    if info.value < 0:
        raise ValueError(f"xtgevc3 backward: info={info.value} (complex or infinite selected eigenvalue)")
    out = (dS, dP) if Z is None else (dS, dP, dZ)
    return out + (info.value,) if return_info else out
