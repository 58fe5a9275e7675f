"""kronmult_amd -- batched Kronecker-product matrix-vector multiply on MI355X.

    Y(:,k) += kron(A1(k), ..., Ad(k)) * X(:,k),   d <= 6

Python host mirror of the reference's entry points (hipdac-lab/kronmult:
kronmult{1..6}_batched.cpp, kronmult{1..6}_xbatched.cpp, the
kronmult{1..6}_vbatched kernels, kgemm_n{n,t}_batched.cpp) over the C ABI
of libkronmult_amd.so (include/kronmult_amd.h).  Argument meaning and order
follow the reference; every data argument is device memory, given as a torch
tensor or as an integer device address.  Pointer arrays (Aarray, Xarray, ...
of the pointer layouts) are device tensors of int64 addresses, as in the
reference where the kernels dereference them on the device.

Errors: the reference does not check arguments; here a non-OK status raises
KronError before anything is launched.  There is no CPU fallback: when the
HIP library is missing every call raises.
"""
import ctypes
import functools
import os

__all__ = [
    "KronError", "lib", "ptr_array", "XbatchedPlan", "MixedPlan", "MixedClass",
    "kronmult_batched", "kronmult_batched_adjoint", "adjoint_workspace_elems",
    "kronmult_xbatched_adjoint",
    "kronmult_xbatched", "kronmult_vbatched",
    "kronmult_vbatched_plain", "kronmult_pbatched",
    "kgemm_nn_batched", "kgemm_nt_batched",
] + ["kronmult%d_%s" % (d, k) for d in range(1, 7)
     for k in ("batched", "pbatched", "xbatched", "vbatched")]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkronmult_amd.so")

KRON_OK, KRON_EINVAL, KRON_EWORKSPACE, KRON_EHIP, KRON_ENOTSUP, KRON_ETOOBIG = range(6)
KRON_NOALLOC_MAX_BATCH = 1048576


class KronError(RuntimeError):
    """A kron_status other than KRON_OK."""

    def __init__(self, fn, status, message):
        super().__init__("%s: %s" % (fn, message))
        self.status = status


_c_int, _c_vp, _c_sz = ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t
_c_ip = ctypes.POINTER(ctypes.c_int)

_SIGS = {
    "kron_batched_%s": [_c_int, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_int, _c_vp],
    "kron_xbatched_%s": [_c_int, _c_int, _c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_int, _c_int, _c_vp],
    "kron_pbatched_%s": [_c_int, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_int, _c_vp],
    "kron_vbatched_%s": [_c_int, _c_ip, _c_ip, _c_vp, _c_vp, _c_vp, _c_vp, _c_sz, _c_int, _c_vp],
    "kron_vbatched_plain_%s": [_c_int, _c_ip, _c_ip, _c_vp, _c_vp, _c_vp, _c_vp, _c_sz, _c_int,
                               _c_vp],
    "kron_batched_adjoint_%s": [_c_int, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_int,
                                _c_vp],
    "kron_xbatched_adjoint_%s": [_c_int, _c_int, _c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_int,
                                 _c_vp, _c_int, _c_vp],
}


@functools.lru_cache(maxsize=1)
def lib():
    """Load libkronmult_amd.so (built in-tree by kronmult_amd.build)."""
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "kronmult_amd: %s is not built (run `python -m kronmult_amd.build`); "
            "there is no fallback path" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    for sfx, real in (("f64", ctypes.c_double), ("f32", ctypes.c_float)):
        for name, args in _SIGS.items():
            f = getattr(L, name % sfx)
            f.argtypes = args
            f.restype = _c_int
        for tr in ("nn", "nt"):
            f = getattr(L, "kron_kgemm_%s_batched_%s" % (tr, sfx))
            f.argtypes = [_c_int, _c_int, _c_int, real, _c_vp, _c_vp, _c_vp, _c_vp, real,
                          _c_vp, _c_vp, _c_int, _c_vp]
            f.restype = _c_int
    for sfx in ("f64", "f32"):
        f = getattr(L, "kron_xbatched_plan_create_" + sfx)
        f.argtypes = [ctypes.POINTER(_c_vp), _c_int, _c_int, _c_vp, _c_int, _c_vp, _c_vp, _c_vp,
                      _c_int, _c_int, _c_vp]
        f.restype = _c_int
    L.kron_xbatched_noalloc_f64.argtypes = _SIGS["kron_xbatched_%s"]
    L.kron_xbatched_noalloc_f64.restype = _c_int
    L.kron_xbatched_plan_execute.argtypes = [_c_vp, _c_vp]
    L.kron_xbatched_plan_execute.restype = _c_int
    L.kron_xbatched_plan_destroy.argtypes = [_c_vp]
    L.kron_xbatched_plan_destroy.restype = _c_int
    for sfx in ("f64", "f32"):
        f = getattr(L, "kron_mixed_plan_create_" + sfx)
        f.argtypes = [ctypes.POINTER(_c_vp), _c_int, _c_vp, ctypes.c_char_p, _c_int, _c_vp]
        f.restype = _c_int
    for name, args in (("execute", [_c_vp, _c_vp]), ("capture", [_c_vp, _c_int]),
                       ("replay", [_c_vp, _c_vp]), ("destroy", [_c_vp]),
                       ("class_info", [_c_vp, _c_int, _c_ip, _c_ip, _c_ip,
                                       ctypes.POINTER(ctypes.c_double)]),
                       ("launch_info", [_c_vp, _c_ip, _c_ip, _c_ip, _c_ip])):
        f = getattr(L, "kron_mixed_plan_" + name)
        f.argtypes = args
        f.restype = _c_int
    L.kron_batched_adjoint_workspace.argtypes = [_c_int, _c_int, _c_int,
                                                 ctypes.POINTER(ctypes.c_size_t)]
    L.kron_batched_adjoint_workspace.restype = _c_int
    L.kron_status_string.argtypes = [_c_int]
    L.kron_status_string.restype = ctypes.c_char_p
    L.kron_last_hip_error.restype = _c_int
    L.kron_version.restype = _c_int
    return L


def _addr(x):
    """Device address of a tensor / int / None."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def _sfx(dtype):
    import torch
    if dtype in (torch.float64, "float64", "f64", "double"):
        return "f64"
    if dtype in (torch.float32, "float32", "f32", "float"):
        return "f32"
    raise TypeError("kronmult_amd supports float64 and float32, got %r" % (dtype,))


def _stream(stream):
    if stream is not None:
        return stream if isinstance(stream, int) else stream.cuda_stream
    import torch
    return torch.cuda.current_stream().cuda_stream


def _check(fn, st):
    if st != KRON_OK:
        L = lib()
        msg = L.kron_status_string(st).decode()
        if st == KRON_EHIP:
            msg += " (hipError %d)" % L.kron_last_hip_error()
        raise KronError(fn, st, msg)


def ptr_array(items, device="cuda"):
    """Device int64 tensor of addresses (tensors or ints) -- a pointer array."""
    import torch
    return torch.tensor([_addr(t) for t in items], dtype=torch.int64, device=device)


def kronmult_batched(d, n, Aarray, Xarray, Yarray, Warray, batchCount, stream=None):
    """kronmult{d}_batched (kronmult6_batched.cpp:5-25).

    Aarray (n,n,d,batch), Xarray/Yarray/Warray (n^d, batch), column-major.
    Y += kron(A1..Ad) X per entry; X is read-only, W scratch."""
    sfx = _sfx(Aarray.dtype)
    f = getattr(lib(), "kron_batched_" + sfx)
    _check("kronmult%d_batched" % d,
           f(d, n, _addr(Aarray), _addr(Xarray), _addr(Yarray), _addr(Warray), batchCount,
             _stream(stream)))


def kronmult_batched_adjoint(d, n, Aarray, Xarray, dYarray, dXarray, dAarray, Warray,
                             batchCount, stream=None):
    """Adjoint of kronmult{d}_batched (no counterpart in the reference):
    dX = kron(A1..Ad)^T dY and dA = d<dY, kron(A) X>/dA, both overwritten.
    Warray: 2 * n^d * batchCount elements of scratch, or None when the entry
    fits on chip."""
    sfx = _sfx(Aarray.dtype)
    f = getattr(lib(), "kron_batched_adjoint_" + sfx)
    _check("kronmult%d_batched_adjoint" % d,
           f(d, n, _addr(Aarray), _addr(Xarray), _addr(dYarray), _addr(dXarray),
             _addr(dAarray), _addr(Warray), batchCount, _stream(stream)))


def kronmult_xbatched_adjoint(d, n, Aarray, lda, Xarray, dYarray, dXarray, dAarray, ldda,
                              Warray, batchCount, dtype=None, stream=None):
    """Adjoint of kronmult{d}_xbatched (no counterpart in the reference), over
    int64 device pointer arrays in the layout of kronmult_xbatched.hpp:79-83:
    dXarray[b] += kron(A)^T dYarray[b] and dAarray[k + b*d] (n x n, leading
    dimension ldda) += the gradient of factor k, with device atomics (entries
    may share any target).  Warray: adjoint_workspace_elems(d, n) *
    batchCount contiguous elements, or None when that is 0."""
    import torch
    sfx = _sfx(dtype if dtype is not None else torch.float64)
    f = getattr(lib(), "kron_xbatched_adjoint_" + sfx)
    _check("kronmult%d_xbatched_adjoint" % d,
           f(d, n, _addr(Aarray), lda, _addr(Xarray), _addr(dYarray), _addr(dXarray),
             _addr(dAarray), ldda, _addr(Warray), batchCount, _stream(stream)))


def adjoint_workspace_elems(d, n, dtype):
    """Elements of W per entry that kronmult_batched_adjoint and
    kronmult_xbatched_adjoint need at (d, n): 0 when the entry fits on chip
    (W may be None), else 2 * n^d."""
    out = ctypes.c_size_t()
    _check("kron_batched_adjoint_workspace",
           lib().kron_batched_adjoint_workspace(d, n, 8 if _sfx(dtype) == "f64" else 4,
                                                ctypes.byref(out)))
    return out.value


def kronmult_pbatched(d, n, Aarray, Xarray, Yarray, Warray, batchCount, dtype=None,
                      stream=None):
    """kronmult{d}_pbatched (kronmult6_pbatched.cpp:5-26): A (n,n,d,batch)
    contiguous, X/Y/W as device arrays of pointers (Y may be shared)."""
    import torch
    sfx = _sfx(dtype if dtype is not None else (
        Aarray.dtype if hasattr(Aarray, "dtype") else torch.float64))
    f = getattr(lib(), "kron_pbatched_" + sfx)
    _check("kronmult%d_pbatched" % d,
           f(d, n, _addr(Aarray), _addr(Xarray), _addr(Yarray), _addr(Warray), batchCount,
             _stream(stream)))


def kronmult_xbatched(d, n, Aarray, lda, Xarray, Yarray, Warray, batchCount,
                      subbatchCount=0, dtype=None, stream=None):
    """kronmult{d}_xbatched (kronmult6_xbatched.cpp:5-39).

    Aarray: int64 device pointers, Aarray[k + b*d] -> n x n factor (ld lda);
    Xarray/Yarray/Warray: int64 device pointers (Y entries may coincide)."""
    import torch
    sfx = _sfx(dtype if dtype is not None else torch.float64)
    f = getattr(lib(), "kron_xbatched_" + sfx)
    _check("kronmult%d_xbatched" % d,
           f(d, n, _addr(Aarray), lda, _addr(Xarray), _addr(Yarray), _addr(Warray),
             batchCount, subbatchCount, _stream(stream)))


def kronmult_vbatched(m, n, Aarray, Xarray, Yarray, W, Wcapacity, batchCount,
                      dtype=None, stream=None):
    """kronmult{d}_vbatched (kronmult6_vbatched.hpp:14-58), d = len(m).

    Factor k is m[k] x n[k] with leading dimension m[k]; W is a scratch buffer
    of Wcapacity bytes."""
    import torch
    d = len(m)
    if len(n) != d:
        raise KronError("kronmult_vbatched", KRON_EINVAL, "len(m) != len(n)")
    sfx = _sfx(dtype if dtype is not None else torch.float64)
    f = getattr(lib(), "kron_vbatched_" + sfx)
    mm = (ctypes.c_int * d)(*m)
    nn = (ctypes.c_int * d)(*n)
    _check("kronmult%d_vbatched" % d,
           f(d, mm, nn, _addr(Aarray), _addr(Xarray), _addr(Yarray), _addr(W), Wcapacity,
             batchCount, _stream(stream)))


def kronmult_vbatched_plain(m, n, Aarray, Xarray, Yarray, W, Wcapacity, batchCount,
                            dtype=None, stream=None):
    """kron_vbatched_plain_*: the kronmult_vbatched update with the reference's
    per-entry structure (one workgroup per entry, one pass through W per mode);
    the baseline the mixed-shape plan is compared with."""
    import torch
    d = len(m)
    sfx = _sfx(dtype if dtype is not None else torch.float64)
    f = getattr(lib(), "kron_vbatched_plain_" + sfx)
    mm = (ctypes.c_int * d)(*m)
    nn = (ctypes.c_int * d)(*n)
    _check("kron_vbatched_plain",
           f(d, mm, nn, _addr(Aarray), _addr(Xarray), _addr(Yarray), _addr(W), Wcapacity,
             batchCount, _stream(stream)))


class XbatchedPlan:
    """kron_xbatched_plan_*: the Y-target order of an xbatched call computed
    once, reused by every execute() while the pointer arrays stay unchanged."""

    def __init__(self, d, n, Aarray, lda, Xarray, Yarray, Warray, batchCount,
                 subbatchCount=0, dtype=None, stream=None):
        import torch
        sfx = _sfx(dtype if dtype is not None else torch.float64)
        self._h = _c_vp()
        self._keep = (Aarray, Xarray, Yarray, Warray)
        f = getattr(lib(), "kron_xbatched_plan_create_" + sfx)
        _check("kron_xbatched_plan_create",
               f(ctypes.byref(self._h), d, n, _addr(Aarray), lda, _addr(Xarray),
                 _addr(Yarray), _addr(Warray), batchCount, subbatchCount, _stream(stream)))

    def execute(self, stream=None):
        _check("kron_xbatched_plan_execute",
               lib().kron_xbatched_plan_execute(self._h, _stream(stream)))

    def close(self):
        if self._h:
            lib().kron_xbatched_plan_destroy(self._h)
            self._h = _c_vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


KRON_MIXED_TUNE, KRON_MIXED_ATOMIC = 0x1, 0x2
KRON_MIXED_FORCE_V0, KRON_MIXED_FORCE_V1, KRON_MIXED_FORCE_V2 = 0x100, 0x200, 0x400
MIXED_KINDS = {0: "launch", 1: "onchip", 2: "packed", 3: "twopass", 4: "slab"}


class _CMixedClass(ctypes.Structure):
    # kron_mixed_class_f64 / _f32 (include/kronmult_amd.h)
    _fields_ = [("d", _c_int), ("n", _c_int), ("lda", _c_int), ("batchCount", _c_int),
                ("Aarray", _c_vp), ("pX", _c_vp), ("pY", _c_vp)]


class MixedClass:
    """One square (d, n) class of a mixed plan: xbatched-style pointer arrays
    (Aarray[k + b*d], pX[b], pY[b]) on the device."""

    def __init__(self, d, n, Aarray, Xarray, Yarray, batchCount, lda=0):
        self.d, self.n, self.lda, self.batchCount = d, n, lda, batchCount
        self.Aarray, self.Xarray, self.Yarray = Aarray, Xarray, Yarray


class MixedPlan:
    """kron_mixed_plan_*: many (d, n) classes applied in one persistent launch.

    tune_cache: path of the autotuner's cache file (or None); tune=True times
    the mappings of classes missing from it.  capture(applies) records that
    many executes in a hipGraph; replay() launches it."""

    def __init__(self, classes, dtype=None, tune_cache=None, tune=False, flags=0, stream=None):
        import torch
        sfx = _sfx(dtype if dtype is not None else torch.float64)
        self._h = _c_vp()
        self._keep = list(classes)
        arr = (_CMixedClass * max(1, len(classes)))()
        for i, c in enumerate(classes):
            arr[i] = _CMixedClass(c.d, c.n, c.lda, c.batchCount, _addr(c.Aarray),
                                  _addr(c.Xarray), _addr(c.Yarray))
        fl = flags | (KRON_MIXED_TUNE if tune else 0)
        path = tune_cache.encode() if tune_cache else None
        f = getattr(lib(), "kron_mixed_plan_create_" + sfx)
        _check("kron_mixed_plan_create",
               f(ctypes.byref(self._h), len(classes), ctypes.cast(arr, _c_vp), path, fl,
                 _stream(stream)))

    def execute(self, stream=None):
        _check("kron_mixed_plan_execute", lib().kron_mixed_plan_execute(self._h, _stream(stream)))

    def capture(self, applies):
        _check("kron_mixed_plan_capture", lib().kron_mixed_plan_capture(self._h, applies))

    def replay(self, stream=None):
        _check("kron_mixed_plan_replay", lib().kron_mixed_plan_replay(self._h, _stream(stream)))

    def class_info(self, cls):
        kind, var, atom = _c_int(), _c_int(), _c_int()
        cost = ctypes.c_double()
        _check("kron_mixed_plan_class_info",
               lib().kron_mixed_plan_class_info(self._h, cls, ctypes.byref(kind),
                                                ctypes.byref(var), ctypes.byref(atom),
                                                ctypes.byref(cost)))
        return {"kind": MIXED_KINDS[kind.value], "variant": var.value, "atomic": atom.value,
                "cost": cost.value}

    def launch_info(self):
        items, grid, lds, nl = _c_int(), _c_int(), _c_int(), _c_int()
        _check("kron_mixed_plan_launch_info",
               lib().kron_mixed_plan_launch_info(self._h, ctypes.byref(items), ctypes.byref(grid),
                                                 ctypes.byref(lds), ctypes.byref(nl)))
        return {"items": items.value, "grid": grid.value, "lds_bytes": lds.value,
                "launches": nl.value}

    def close(self):
        if self._h:
            lib().kron_mixed_plan_destroy(self._h)
            self._h = _c_vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _kgemm(tr, mm, nn, kk, alpha, Aarray, ldAarray, Barray, ldBarray, beta, Carray,
           ldCarray, batchCount, dtype, stream):
    import torch
    sfx = _sfx(dtype if dtype is not None else torch.float64)
    f = getattr(lib(), "kron_kgemm_%s_batched_%s" % (tr, sfx))
    _check("kgemm_%s_batched" % tr,
           f(mm, nn, kk, alpha, _addr(Aarray), _addr(ldAarray), _addr(Barray),
             _addr(ldBarray), beta, _addr(Carray), _addr(ldCarray), batchCount,
             _stream(stream)))


def kgemm_nn_batched(mm, nn, kk, alpha, Aarray, ldAarray, Barray, ldBarray, beta, Carray,
                     ldCarray, batchCount, dtype=None, stream=None):
    """kgemm_nn_batched (kgemm_nn_batched.cpp:9-42): C_b = alpha A_b B_b + beta C_b."""
    _kgemm("nn", mm, nn, kk, alpha, Aarray, ldAarray, Barray, ldBarray, beta, Carray,
           ldCarray, batchCount, dtype, stream)


def kgemm_nt_batched(mm, nn, kk, alpha, Aarray, ldAarray, Barray, ldBarray, beta, Carray,
                     ldCarray, batchCount, dtype=None, stream=None):
    """kgemm_nt_batched (kgemm_nt_batched.cpp:8-41): C_b = alpha A_b B_b^T + beta C_b."""
    _kgemm("nt", mm, nn, kk, alpha, Aarray, ldAarray, Barray, ldBarray, beta, Carray,
           ldCarray, batchCount, dtype, stream)


def _bind(d):
    def batched(n, Aarray, Xarray, Yarray, Warray, batchCount, stream=None):
        return kronmult_batched(d, n, Aarray, Xarray, Yarray, Warray, batchCount, stream)

    def xbatched(n, Aarray, lda, Xarray, Yarray, Warray, batchCount, subbatchCount=0,
                 dtype=None, stream=None):
        return kronmult_xbatched(d, n, Aarray, lda, Xarray, Yarray, Warray, batchCount,
                                 subbatchCount, dtype, stream)

    def vbatched(*args, dtype=None, stream=None):
        # (m1, n1, ..., md, nd, Aarray, pX, pY, W, Wcapacity, batchCount)
        if len(args) != 2 * d + 6:
            raise TypeError("kronmult%d_vbatched takes %d positional arguments" % (d, 2 * d + 6))
        mn = args[:2 * d]
        rest = args[2 * d:]
        return kronmult_vbatched(list(mn[0::2]), list(mn[1::2]), *rest, dtype=dtype,
                                 stream=stream)

    def pbatched(n, Aarray, Xarray, Yarray, Warray, batchCount, dtype=None, stream=None):
        return kronmult_pbatched(d, n, Aarray, Xarray, Yarray, Warray, batchCount, dtype,
                                 stream)

    for f, kind in ((batched, "batched"), (xbatched, "xbatched"), (vbatched, "vbatched"),
                    (pbatched, "pbatched")):
        f.__name__ = "kronmult%d_%s" % (d, kind)
        f.__doc__ = "kronmult%d_%s -- see kronmult_%s." % (d, kind, kind)
    return batched, xbatched, vbatched, pbatched


for _d in range(1, 7):
    (globals()["kronmult%d_batched" % _d], globals()["kronmult%d_xbatched" % _d],
     globals()["kronmult%d_vbatched" % _d], globals()["kronmult%d_pbatched" % _d]) = _bind(_d)
del _d
