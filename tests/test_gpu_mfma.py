"""The MFMA large-n path (n in {16, 32}, batched layout) against the oracle:
integer-valued inputs bit-exact (every partial sum is an exact integer, so the
matrix cores' summation order cannot change the result), normal(0, 1/n)
inputs at 1e-12 relative in FP64, and FP32 against the FP64 oracle at the
tolerance stated per test."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import kronmult_amd as K  # noqa: E402
import oracle  # noqa: E402
from gpu_util import dev, host  # noqa: E402

O = oracle.oracle_lib()

# every (d, n) with an MFMA kernel: one pass (n^d <= 8192) or two passes
MFMA_SHAPES = [(2, 16, 6), (3, 16, 5), (4, 16, 3), (5, 16, 2), (2, 32, 5), (3, 32, 4)]


def _oracle(d, n, A, X, Y0):
    Y = Y0.astype(np.float64)
    W = np.zeros(X.shape)
    oracle.batched(O, d, n, np.ascontiguousarray(A, dtype=np.float64),
                   X.astype(np.float64), Y, W, X.shape[0])
    return Y


def _gpu(d, n, A, X, Y0):
    Ad, Xd, Yd = dev(A), dev(X), dev(Y0)
    Wd = torch.full_like(Xd, float("nan"))
    X_before = Xd.clone()
    K.kronmult_batched(d, n, Ad, Xd, Yd, Wd, X.shape[0])
    out = host(Yd)
    assert torch.equal(Xd, X_before), "X must not be written"
    return out


def _normal_case(d, n, batch, seed, dtype=np.float64):
    """A and X i.i.d. normal(0, 1/n) (Box-Muller on splitmix64)."""
    s = 1.0 / np.sqrt(n)
    A = H.normal(batch * d * n * n, H.sub_seed(seed, 1), s).reshape(batch, -1).astype(dtype)
    X = H.normal(batch * n ** d, H.sub_seed(seed, 2), s).reshape(batch, -1).astype(dtype)
    Y0 = H.uniform(batch * n ** d, H.sub_seed(seed, 3)).reshape(batch, -1).astype(dtype)
    return A, X, Y0


@pytest.mark.parametrize("d,n,batch", MFMA_SHAPES)
def test_mfma_integer_bit_exact(d, n, batch):
    A = H.integers(batch * d * n * n, 7 + d, -1, 1).reshape(batch, -1)
    X = H.integers(batch * n ** d, 8 + d, -2, 2).reshape(batch, -1)
    Y0 = H.integers(batch * n ** d, 9 + d, -4, 4).reshape(batch, -1)
    assert np.array_equal(_gpu(d, n, A, X, Y0), _oracle(d, n, A, X, Y0))


@pytest.mark.parametrize("d,n,batch", MFMA_SHAPES)
def test_mfma_f64_normal(d, n, batch):
    A, X, Y0 = _normal_case(d, n, batch, 100 * d + n)
    assert H.rel_err(_gpu(d, n, A, X, Y0), _oracle(d, n, A, X, Y0)) < 1e-12


@pytest.mark.parametrize("d,n,batch", MFMA_SHAPES)
def test_mfma_f32_vs_f64_oracle(d, n, batch):
    """FP32 inputs and arithmetic against the FP64 oracle on the same (FP32-
    rounded) inputs: relative error below 2e-6 (a few float ulps for sums of
    up to n*d terms of normal(0, 1/n) products)."""
    A, X, Y0 = _normal_case(d, n, batch, 200 * d + n, np.float32)
    got = _gpu(d, n, A, X, Y0)
    want = _oracle(d, n, A.astype(np.float64), X.astype(np.float64), Y0.astype(np.float64))
    assert H.rel_err(got, want) < 2e-6


def test_mfma_needs_workspace():
    """Two-pass shapes go through W, as the reference's batched layout does."""
    d, n = 3, 32
    A = torch.zeros(d * n * n, dtype=torch.float64, device="cuda:0")
    X = torch.zeros(n ** d, dtype=torch.float64, device="cuda:0")
    with pytest.raises(K.KronError) as e:
        K.kronmult_batched(d, n, A, X, X.clone(), None, 1)
    assert e.value.status == K.KRON_EWORKSPACE


def test_mfma_repeated_calls_accumulate():
    d, n, batch = 3, 32, 3
    A, X, Y0 = _normal_case(d, n, batch, 77)
    Ad, Xd, Yd = dev(A), dev(X), dev(Y0)
    Wd = torch.empty_like(Xd)
    for _ in range(3):
        K.kronmult_batched(d, n, Ad, Xd, Yd, Wd, batch)
    Y = Y0.copy()
    for _ in range(3):
        Y = _oracle(d, n, A, X, Y)
    assert H.rel_err(host(Yd), Y) < 1e-12
