"""Parity at the workload sizes the library is quoted on, and runs of every
kernel family's persistent multi-entry walk.

* d=6 n=4 batched, 1,024 entries, uniform inputs (the reference's batched
  CPU case, kronmult6_batched.hpp:14-94): FP64 within 1e-12 of the oracle.
* d=3 n=32 and d=4 n=16, 2,048 entries each, normal(0, 1/n) inputs on the
  matrix cores: FP64 within 1e-12 of the oracle; FP32 against the FP64 oracle.
* d=6 n=6, 4,096 entries, forward and adjoint through the autograd wrapper:
  the forward against the oracle for every entry; the inner-product
  identities <dY, K X> = <K^T dY, X> = <A_k, dA_k> for every entry and every
  k; dX and dA against the per-mode formula on a spread sample of entries.
* Integer-valued inputs (exact in FP64 whatever the summation order) with
  more entries than the kernel keeps resident and not a multiple of it, in
  the batched and the ungrouped pointer layouts, for the plane, fused,
  generic (single- and two-pass), matrix-core and adjoint kernels.
"""
import numpy as np
import pytest

import helpers as H
import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import kronmult_amd as K  # noqa: E402
from kronmult_amd.torch_ops import kron_matvec  # noqa: E402
from gpu_util import DEV, dev, host, ptrs  # noqa: E402

O = oracle.oracle_lib()
F64_TOL = 1e-12


def _oracle_batched(d, n, A, X, Y0):
    Y, Xc, W = Y0.astype(np.float64), X.astype(np.float64), np.zeros(X.shape)
    oracle.batched(O, d, n, np.ascontiguousarray(A, dtype=np.float64), Xc, Y, W, X.shape[0])
    return Y


def _gpu_batched(d, n, A, X, Y0):
    Ad, Xd, Yd = dev(A), dev(X), dev(Y0)
    Wd = torch.full_like(Xd, float("nan"))
    K.kronmult_batched(d, n, Ad, Xd, Yd, Wd, X.shape[0])
    out = host(Yd)
    assert torch.equal(Xd, dev(X)), "X must not be written"
    return out


def _int_batched(d, n, batch, seed, alo=-2, ahi=2, xlo=-8, xhi=8):
    E = n ** d
    A = H.integers(batch * d * n * n, H.sub_seed(seed, 1), alo, ahi).reshape(batch, d, n * n)
    X = H.integers(batch * E, H.sub_seed(seed, 2), xlo, xhi).reshape(batch, E)
    Y0 = H.integers(batch * E, H.sub_seed(seed, 3), -9, 9).reshape(batch, E)
    return A, X, Y0


# --- the quoted workloads at their stated sizes -------------------------------

def test_cfg_batched_d6n4_1024_uniform():
    """kronmult6_batched, d=6 n=4, 1,024 entries, A and X uniform [-1,1],
    Y zero, every entry its own Y (test_kronmult6_batched.cpp:511-567 runs
    the same layout up to 128 entries)."""
    d, n, batch = 6, 4, 1024
    E = n ** d
    A = H.uniform(batch * d * n * n, 2024).reshape(batch, d, n * n)
    X = H.uniform(batch * E, 2025).reshape(batch, E)
    Y0 = np.zeros((batch, E))
    got = _gpu_batched(d, n, A, X, Y0)
    want = _oracle_batched(d, n, A, X, Y0)
    assert H.rel_err(got, want) < F64_TOL
    # the dense-Kronecker brute force on a few entries pins the oracle itself
    for b in (0, 511, 1023):
        f = [H.factor_matrix(A[b, k], n, n, n) for k in range(d)]
        assert H.rel_err(want[b], H.kron_numpy(f, X[b])) < 1e-13


def _normal_case(d, n, batch, seed, dtype=np.float64):
    s = 1.0 / np.sqrt(n)
    A = H.normal(batch * d * n * n, H.sub_seed(seed, 1), s).reshape(batch, -1).astype(dtype)
    X = H.normal(batch * n ** d, H.sub_seed(seed, 2), s).reshape(batch, -1).astype(dtype)
    return A, X


@pytest.mark.parametrize("d,n", [(3, 32), (4, 16)])
def test_cfg_large_n_2048_f64(d, n):
    """Large n on the matrix cores, 2,048 entries, A and X normal(0, 1/n)."""
    batch = 2048
    A, X = _normal_case(d, n, batch, 500 + 10 * d + n)
    Y0 = np.zeros_like(X)
    got = _gpu_batched(d, n, A, X, Y0)
    want = _oracle_batched(d, n, A, X, Y0)
    assert H.rel_err(got, want) < F64_TOL


@pytest.mark.parametrize("d,n", [(3, 32), (4, 16)])
def test_cfg_large_n_2048_f32(d, n):
    """FP32 data and arithmetic on the matrix cores against the FP64 oracle
    on the same FP32-rounded inputs.  The error stays within a few float
    ulps of the output scale: 2e-6 in the relative measure."""
    batch = 2048
    A, X = _normal_case(d, n, batch, 600 + 10 * d + n, np.float32)
    Y0 = np.zeros_like(X)
    got = _gpu_batched(d, n, A, X, Y0)
    want = _oracle_batched(d, n, A.astype(np.float64), X.astype(np.float64),
                           Y0.astype(np.float64))
    err = H.rel_err(got, want)
    assert err < 2e-6, err
    # stated error relative to the output's scale
    assert H.norm_rel_err(got, want) < 2e-6


def _rowdot(a, b):
    return np.einsum("ij,ij->i", a, b)


def test_cfg_fwd_adjoint_d6n6_4096():
    """Forward and adjoint at d=6 n=6, 4,096 entries, through kron_matvec."""
    d, n, batch = 6, 6, 4096
    E = n ** d
    A = H.uniform(batch * d * n * n, 4001).reshape(batch, d, n * n)
    X = H.uniform(batch * E, 4002).reshape(batch, E)
    dY = H.uniform(batch * E, 4003).reshape(batch, E)
    # kron_matvec takes A[b, k] as a (row, column) matrix
    Am = torch.from_numpy(A.reshape(batch, d, n, n)).transpose(-1, -2).contiguous()
    Ag = Am.to(DEV).requires_grad_(True)
    Xg = torch.from_numpy(X).to(DEV).requires_grad_(True)
    Y = kron_matvec(Ag, Xg)
    Y.backward(torch.from_numpy(dY).to(DEV))
    Yh = host(Y.detach())
    dX = host(Xg.grad)
    dA = host(Ag.grad.transpose(-1, -2).contiguous()).reshape(batch, d, n * n)
    del Y, Ag, Xg
    torch.cuda.empty_cache()

    # forward: every entry against the oracle
    want = _oracle_batched(d, n, A, X, np.zeros_like(X))
    assert H.rel_err(Yh, want) < F64_TOL

    # inner-product identities for every entry (and every factor)
    L = _rowdot(dY, Yh)
    scale = _rowdot(np.abs(dY), np.abs(Yh)) + 1.0
    assert np.max(np.abs(L - _rowdot(dX, X)) / scale) < F64_TOL
    for k in range(d):
        # L is linear in A_k, so <A_k, dL/dA_k> = L
        assert np.max(np.abs(L - _rowdot(A[:, k], dA[:, k])) / scale) < F64_TOL, k

    # dX and dA against the per-mode formula on a spread sample
    idx = np.unique(np.linspace(0, batch - 1, 32).astype(np.int64))
    rdX, rdA = H.grads_batched(d, n, A[idx], X[idx], dY[idx])
    assert H.rel_err(dX[idx], rdX) < F64_TOL
    assert H.rel_err(dA[idx], rdA) < F64_TOL


# --- persistent walks: more entries than resident workgroups -------------------

# plane (6,4); fused (3,6) 4 entries per workgroup, (5,5), (5,6) 512 threads,
# (6,5) 1,024 threads, (4,8), (2,8); generic single pass (3,12), (4,9) two
# passes; matrix cores (3,16) one pass, (3,32) two passes.
WALK_BATCHED = [(6, 4, 3331), (3, 6, 20011), (5, 5, 3001), (5, 6, 1201), (6, 5, 1103),
                (4, 8, 2111), (2, 8, 9001), (3, 12, 3331), (4, 9, 1601), (3, 16, 3331),
                (3, 32, 1031)]


@pytest.mark.parametrize("d,n,batch", WALK_BATCHED)
def test_walk_batched_integer(d, n, batch):
    lo = -1 if n >= 9 else -2
    A, X, Y0 = _int_batched(d, n, batch, 100 * d + n, lo, -lo, -4, 4)
    assert np.array_equal(_gpu_batched(d, n, A, X, Y0), _oracle_batched(d, n, A, X, Y0))


def _ptr_case(d, n, batch, nx, ny, lda, seed):
    E = n ** d
    A = H.integers(batch * d * n * n, H.sub_seed(seed, 1), -1, 1).reshape(batch * d, n, n)
    A_pool = np.full((batch * d, n, lda), np.nan)
    A_pool[:, :, :n] = A
    A_pool = A_pool.reshape(batch * d, lda * n)
    X_pool = H.integers(nx * E, H.sub_seed(seed, 2), -4, 4).reshape(nx, E)
    x_map = (H.splitmix64(batch, H.sub_seed(seed, 3)) % np.uint64(nx)).astype(np.int64)
    y_map = (H.splitmix64(batch, H.sub_seed(seed, 4)) % np.uint64(ny)).astype(np.int64)
    return A_pool, X_pool, x_map, y_map


# ungrouped pointer layout: fewer than 2,048 entries (no Y-target sort), more
# than the resident workgroups of each kernel
WALK_PTR = [(6, 4, 1999, 5), (5, 6, 1201, 6), (5, 7, 1201, 8), (6, 5, 1103, 7), (4, 8, 1777, 4),
            (3, 12, 1999, 5), (4, 9, 1601, 9), (3, 16, 1999, 3), (3, 32, 1031, 2)]


@pytest.mark.parametrize("d,n,batch,ny", WALK_PTR)
def test_walk_xbatched_integer(d, n, batch, ny):
    lda = n + 2
    A_pool, X_pool, x_map, y_map = _ptr_case(d, n, batch, 37, ny, lda, 300 * d + n)
    E = n ** d
    Y0 = np.zeros((ny, E))
    Ad, Xd, Yd = dev(A_pool), dev(X_pool), dev(Y0)
    nslot = 2 * batch
    Wd = torch.full((nslot, E), float("nan"), dtype=torch.float64, device=DEV)
    K.kronmult_xbatched(d, n, ptrs(Ad, np.arange(batch * d)), lda, ptrs(Xd, x_map),
                        ptrs(Yd, y_map), ptrs(Wd, np.arange(nslot)), batch, batch)
    got = host(Yd)
    Yw = Y0.copy()
    oracle.xbatched(O, d, n, A_pool, np.arange(batch * d).reshape(batch, d), lda, X_pool,
                    x_map, Yw, y_map, np.zeros((nslot, E)), np.arange(nslot), batch)
    assert np.array_equal(got, Yw)


@pytest.mark.parametrize("d,n", [(5, 6), (5, 7)])
def test_xbatched_shared_y_integer_wide_lda(d, n):
    """Shared Y, lda = n + 2, 300 entries (ungrouped): the 512-thread fused
    geometries through the pointer layout."""
    batch, ny, lda = 300, 9, n + 2
    A_pool, X_pool, x_map, y_map = _ptr_case(d, n, batch, 40, ny, lda, 77 + n)
    E = n ** d
    Y0 = np.zeros((ny, E))
    Ad, Xd, Yd = dev(A_pool), dev(X_pool), dev(Y0)
    Wd = torch.full((2 * batch, E), float("nan"), dtype=torch.float64, device=DEV)
    K.kronmult_xbatched(d, n, ptrs(Ad, np.arange(batch * d)), lda, ptrs(Xd, x_map),
                        ptrs(Yd, y_map), ptrs(Wd, np.arange(2 * batch)), batch)
    Yw = Y0.copy()
    oracle.xbatched(O, d, n, A_pool, np.arange(batch * d).reshape(batch, d), lda, X_pool,
                    x_map, Yw, y_map, np.zeros((2 * batch, E)), np.arange(2 * batch), batch)
    assert np.array_equal(host(Yd), Yw)


# --- the adjoint's multi-item walk ----------------------------------------------

def _gpu_adjoint(d, n, A, X, dY):
    batch, E = X.shape
    Ad, Xd, dYd = dev(A), dev(X), dev(dY)
    dXd = torch.full_like(Xd, float("nan"))
    dAd = torch.full_like(Ad, float("nan"))
    Wd = torch.empty(2 * E * batch, dtype=Xd.dtype, device=DEV)
    K.kronmult_batched_adjoint(d, n, Ad, Xd, dYd, dXd, dAd, Wd, batch)
    return host(dXd), host(dAd)


@pytest.mark.parametrize("d,n,batch", [(6, 6, 211), (4, 4, 9001), (5, 7, 150), (3, 5, 20011),
                                       (6, 4, 3331)])
def test_adjoint_walk_integer(d, n, batch):
    """Many items per workgroup (entries x chunks above twice the resident
    grid): integer inputs, so dX and dA are exact whatever the order."""
    A, X, dY = _int_batched(d, n, batch, 900 + 10 * d + n, -1, 1, -3, 3)
    dX, dA = _gpu_adjoint(d, n, A, X, dY)
    rdX, rdA = H.grads_batched(d, n, A, X, dY)
    assert np.array_equal(dX, rdX)
    assert np.array_equal(dA.reshape(rdA.shape), rdA)


@pytest.mark.parametrize("d,n,batch", [(6, 6, 96), (4, 4, 4099), (6, 4, 2100)])
def test_adjoint_walk_f32(d, n, batch):
    """FP32 adjoint with many items per workgroup, against the FP64 formula on
    the same FP32 inputs."""
    A = H.uniform(batch * d * n * n, 31).reshape(batch, d, n * n).astype(np.float32)
    X = H.uniform(batch * n ** d, 32).reshape(batch, -1).astype(np.float32)
    dY = H.uniform(batch * n ** d, 33).reshape(batch, -1).astype(np.float32)
    dX, dA = _gpu_adjoint(d, n, A, X, dY)
    rdX, rdA = H.grads_batched(d, n, A, X, dY)
    # sums of n^(d-1) products for dA: relative to the gradients' scale
    assert H.norm_rel_err(dX, rdX) < 1e-5
    assert H.norm_rel_err(dA.reshape(rdA.shape), rdA) < 1e-5
