"""Oracle for the batched Kronecker matvec -- TEST INFRASTRUCTURE ONLY.

ctypes front end over two CPU libraries:

* ``liboracle.so``: this repository's C restatement of the reference CPU path
  (kron_oracle.c, which cites the reference file:line it follows);
* ``_ref/libkron_ref.so``: the reference's own CPU/OpenMP path compiled from
  its sources (build.py), present only where the reference was available at
  build time.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this package, to check results against; the product (kronmult_amd) never does.

Arrays are numpy, column-major by the reference's convention: a batch of
vectors is a C-contiguous (batch, size) array, i.e. (size, batch) in
Fortran order; factors are (lda * n) slabs.  Pointer layouts are given as
pools plus index maps and turned into host pointer arrays here.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libkron_ref.so")

_vp = ctypes.c_void_p
_i = ctypes.c_int


def _declare(L, prefix, with_omp):
    omp = [_i] if with_omp else []
    for sfx, real in (("f64", ctypes.c_double), ("f32", ctypes.c_float)):
        f = getattr(L, "%s_batched_%s" % (prefix, sfx))
        f.argtypes = [_i, _i, _vp, _vp, _vp, _vp, _i] + omp
        f.restype = _i
        f = getattr(L, "%s_xbatched_%s" % (prefix, sfx))
        f.argtypes = [_i, _i, _vp, _i, _vp, _vp, _vp, _i, _i] + omp
        f.restype = _i
        f = getattr(L, "%s_vbatched_%s" % (prefix, sfx))
        f.argtypes = [_i, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_size_t, _i] + omp
        f.restype = _i
        f = getattr(L, "%s_kgemm_batched_%s" % (prefix, sfx))
        f.argtypes = [ctypes.c_char, _i, _i, _i, real, _vp, _vp, _vp, _vp, real, _vp, _vp, _i]
        f.restype = _i


_cache = {}


def oracle_lib():
    """The C restatement (built on demand with gcc when missing)."""
    if "oracle" not in _cache:
        if not os.path.exists(ORACLE_SO):
            from . import build
            build.build_oracle()
        L = ctypes.CDLL(ORACLE_SO)
        _declare(L, "oracle", True)
        L.oracle_dense_kron_f64.argtypes = [_i, _vp, _vp, _vp, _vp, _vp, _vp]
        L.oracle_dense_kron_f64.restype = None
        L.oracle_fill_uniform_f64.argtypes = [_vp, ctypes.c_size_t, ctypes.c_uint64]
        L.oracle_fill_uniform_f64.restype = None
        L.oracle_set_threads.argtypes = [_i]
        L.oracle_set_threads.restype = None
        _cache["oracle"] = L
    return _cache["oracle"]


def ref_lib():
    """The compiled reference CPU path, or None where it was not built."""
    if "ref" not in _cache:
        L = None
        if os.path.exists(REF_SO):
            L = ctypes.CDLL(REF_SO)
            _declare(L, "ref", False)
            L.ref_pbatched_f64.argtypes = [_i, _i, _vp, _vp, _vp, _vp, _i]
            L.ref_pbatched_f64.restype = _i
            L.ref_set_threads.argtypes = [_i]
            L.ref_set_threads.restype = None
        _cache["ref"] = L
    return _cache["ref"]


def set_threads(L, threads):
    """OpenMP thread count of an oracle library's loops (C restatement or
    the compiled reference)."""
    if L is oracle_lib():
        L.oracle_set_threads(int(threads))
    else:
        L.ref_set_threads(int(threads))


def _sfx(a):
    if a.dtype == np.float64:
        return "f64"
    if a.dtype == np.float32:
        return "f32"
    raise TypeError(a.dtype)


def _p(a):
    assert a.flags.c_contiguous
    return a.ctypes.data


def _ptrs(pool, idx):
    """Host pointer array to rows pool[idx[i]] (pool is C-contiguous 2-D)."""
    base = pool.ctypes.data
    row = pool.strides[0]
    arr = (ctypes.c_void_p * max(1, len(idx)))()
    for i, r in enumerate(idx):
        arr[i] = base + int(r) * row
    return arr


def _omp(L, use_omp):
    return [int(use_omp)] if L is _cache.get("oracle") else []


def batched(L, d, n, A, X, Y, W, batch, use_omp=True):
    """kronmult{d}_batched: A (batch, d, n*n), X/Y/W (batch, n^d); X clobbered."""
    f = getattr(L, "%s_batched_%s" % ("oracle" if L is oracle_lib() else "ref", _sfx(A)))
    st = f(d, n, _p(A), _p(X), _p(Y), _p(W), batch, *_omp(L, use_omp))
    assert st == 0, st


def xbatched(L, d, n, A_pool, a_map, lda, X_pool, x_map, Y_pool, y_map, W_pool, w_map,
             batch, subbatch=0, use_omp=True):
    """kronmult{d}_xbatched over pools: Aarray[k + b*d] = A_pool[a_map[b, k]], etc.

    The reference copies each X into a pW slot first (kronmult_xbatched.hpp:
    128-133), so pX rows are read-only unless batch == 1."""
    pre = "oracle" if L is oracle_lib() else "ref"
    f = getattr(L, "%s_xbatched_%s" % (pre, _sfx(A_pool)))
    pa = _ptrs(A_pool, np.asarray(a_map).reshape(-1))
    px, py, pw = _ptrs(X_pool, x_map), _ptrs(Y_pool, y_map), _ptrs(W_pool, w_map)
    st = f(d, n, pa, lda, px, py, pw, batch, subbatch, *_omp(L, use_omp))
    assert st == 0, st


def pbatched(L, d, n, A, X_pool, x_map, Y_pool, y_map, batch, use_omp=True):
    """kronmult{d}_pbatched (kronmult6_pbatched.hpp:14-81): A (batch, d, n*n)
    contiguous, X and Y rows of pools through maps, one W vector per entry.
    The reference runs its own pbatched (d = 4 excepted: that file does not
    compile as shipped); the restatement runs the same per-entry recursion
    through its xbatched (factor k of entry b is row b*d + k of A), whose only
    extra step, copying X first, does not change any value."""
    E = X_pool.shape[1]
    if L is not oracle_lib():
        W = np.zeros((batch, E), dtype=X_pool.dtype)
        Xc = X_pool.copy()  # the reference may clobber X
        st = L.ref_pbatched_f64(d, n, _p(np.ascontiguousarray(A)), _ptrs(Xc, x_map),
                                _ptrs(Y_pool, y_map), _ptrs(W, np.arange(batch)), batch)
        assert st == 0, st
        return
    A2 = np.ascontiguousarray(A).reshape(batch * d, -1)
    W = np.zeros((max(2, 2 * batch), E), dtype=X_pool.dtype)
    xbatched(L, d, n, A2, np.arange(batch * d).reshape(batch, d), n, X_pool, x_map, Y_pool,
             y_map, W, np.arange(W.shape[0]), batch, 0, use_omp)


def vbatched(L, m, n, A_pool, a_map, X_pool, x_map, Y_pool, y_map, W, Wcap_bytes, batch,
             use_omp=True):
    """kronmult{d}_vbatched: factor k of entry b is A_pool[a_map[b, k]] (m_k x n_k, ld m_k)."""
    pre = "oracle" if L is oracle_lib() else "ref"
    f = getattr(L, "%s_vbatched_%s" % (pre, _sfx(A_pool)))
    d = len(m)
    mm = (ctypes.c_int * 6)(*m)
    nn = (ctypes.c_int * 6)(*n)
    pa = _ptrs(A_pool, np.asarray(a_map).reshape(-1))
    st = f(d, mm, nn, pa, _ptrs(X_pool, x_map), _ptrs(Y_pool, y_map), _p(W), Wcap_bytes, batch,
           *_omp(L, use_omp))
    assert st == 0, st


def kgemm(L, trans, mm, nn, kk, alpha, A_pool, ldA, B_pool, ldB, beta, C_pool, ldC, batch):
    """kgemm_n{n,t}_batched with entry b using rows b of the pools."""
    pre = "oracle" if L is oracle_lib() else "ref"
    f = getattr(L, "%s_kgemm_batched_%s" % (pre, _sfx(A_pool)))
    idx = np.arange(batch)
    lda = np.ascontiguousarray(ldA, dtype=np.int32)
    ldb = np.ascontiguousarray(ldB, dtype=np.int32)
    ldc = np.ascontiguousarray(ldC, dtype=np.int32)
    st = f(trans.encode(), mm, nn, kk, alpha, _ptrs(A_pool, idx), _p(lda), _ptrs(B_pool, idx),
           _p(ldb), beta, _ptrs(C_pool, idx), _p(ldc), batch)
    assert st == 0, st


def dense_kron(m, n, factors, X):
    """Textbook dense Kronecker brute force in C (double accumulate); returns Y.
    factors[k] is an m[k] x n[k] matrix (any numpy layout)."""
    L = oracle_lib()
    d = len(m)
    mm = (ctypes.c_int * 6)(*m)
    nn = (ctypes.c_int * 6)(*n)
    ld = (ctypes.c_int * 6)(*m)
    # column-major storage with leading dimension m[k]
    fac = [np.ascontiguousarray(np.asarray(f, dtype=np.float64).T) for f in factors]
    pa = (ctypes.c_void_p * 6)(*[f.ctypes.data for f in fac])
    Xc = np.ascontiguousarray(X, dtype=np.float64)
    Y = np.zeros(int(np.prod(m)), dtype=np.float64)
    L.oracle_dense_kron_f64(d, mm, nn, pa, ld, _p(Xc), _p(Y))
    return Y


def fill_uniform(count, seed):
    """splitmix64 uniform [-1, 1) in C (used to pin the numpy generator)."""
    out = np.empty(count, dtype=np.float64)
    oracle_lib().oracle_fill_uniform_f64(out.ctypes.data, count, seed)
    return out
