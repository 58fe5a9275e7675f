"""The FP32 entry points' accuracy contract, measured beside the reference's
own float arithmetic.

The reference's float templates sum every small GEMM in double and round
each mode's result to float (kgemm_nn.hpp:688-701, kgemm_nt.hpp:557-571);
the oracle restates that bit for bit (tests/test_oracle_vs_ref.py pins its
FP32 path to the compiled reference).  This library's FP32 kernels sum each
mode's n terms in FP32 (MFMA f32 accumulation for n >= 16), and merged Y
runs add up to 64 entries in FP32 before the atomic.  On the same float
inputs, both are compared with the FP64 result; the contract is that this
library's error stays within a small multiple of the reference's own float
error (and below the FP32 tolerance of the reference's tests, 5e-4 relative
in norm there, kgemm tests; 1e-5 here)."""
import numpy as np
import pytest

import helpers as H
import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import kronmult_amd as K  # noqa: E402
from gpu_util import DEV, dev, host, ptrs  # noqa: E402

O = oracle.oracle_lib()
U32 = 2.0 ** -24
# this library's FP32 error over the reference's float error, at most
RATIO = 8.0


def _err(got, want):
    return H.norm_rel_err(got, want)


@pytest.mark.parametrize("d,n,batch", [(6, 4, 300), (5, 5, 200), (3, 8, 500), (4, 6, 300),
                                       (2, 7, 1000), (6, 6, 6), (3, 16, 24), (3, 32, 6)])
def test_fp32_batched_error_vs_reference_float(d, n, batch, record_property):
    E = n ** d
    A = H.uniform(batch * d * n * n, 700 + d * 10 + n).reshape(batch, d, n * n).astype(np.float32)
    X = H.uniform(batch * E, 800 + d * 10 + n).reshape(batch, E).astype(np.float32)
    want = np.zeros((batch, E))
    oracle.batched(O, d, n, A.astype(np.float64), X.astype(np.float64), want, np.zeros((batch, E)),
                   batch)
    ref32 = np.zeros((batch, E), dtype=np.float32)
    oracle.batched(O, d, n, A.copy(), X.copy(), ref32, np.zeros((batch, E), dtype=np.float32),
                   batch)
    Xd, Yd = dev(X), torch.zeros((batch, E), dtype=torch.float32, device=DEV)
    K.kronmult_batched(d, n, dev(A), Xd, Yd, torch.empty_like(Xd), batch)
    ours = host(Yd)
    e_ours, e_ref = _err(ours, want), _err(ref32, want)
    record_property("err_ours", e_ours)
    record_property("err_ref", e_ref)
    print("d%d n%d: FP32 error %.3g (reference float %.3g)" % (d, n, e_ours, e_ref))
    assert e_ours <= RATIO * max(e_ref, U32)
    assert e_ours < 1e-5


def test_fp32_xbatched_merged_runs_error_vs_reference_float(record_property):
    """Pointer layout with about 32 entries per Y target (runs merged on the
    device, then added atomically): d = 6, n = 4, 4,096 entries."""
    d, n, batch, nx, ny = 6, 4, 4096, 512, 128
    E = n ** d
    A = H.uniform(batch * d * n * n, 901).reshape(batch * d, n * n).astype(np.float32)
    X = H.uniform(nx * E, 902).reshape(nx, E).astype(np.float32)
    x_map = (H.splitmix64(batch, 903) % np.uint64(nx)).astype(np.int64)
    y_map = (H.splitmix64(batch, 904) % np.uint64(ny)).astype(np.int64)
    a_map = np.arange(batch * d).reshape(batch, d)
    want = np.zeros((ny, E))
    W = np.zeros((2, E))
    oracle.xbatched(O, d, n, A.astype(np.float64), a_map, n, X.astype(np.float64), x_map, want,
                    y_map, W, np.arange(2), batch, 1)
    ref32 = np.zeros((ny, E), dtype=np.float32)
    oracle.xbatched(O, d, n, A, a_map, n, X, x_map, ref32, y_map, W.astype(np.float32),
                    np.arange(2), batch, 1)
    Ad, Xd = dev(A), dev(X)
    Yd = torch.zeros((ny, E), dtype=torch.float32, device=DEV)
    Wd = torch.empty((1, E), dtype=torch.float32, device=DEV)
    K.kronmult_xbatched(d, n, ptrs(Ad, np.arange(batch * d)), n, ptrs(Xd, x_map),
                        ptrs(Yd, y_map), ptrs(Wd, np.zeros(batch, dtype=np.int64)), batch,
                        dtype=torch.float32)
    ours = host(Yd)
    e_ours, e_ref = _err(ours, want), _err(ref32, want)
    record_property("err_ours", e_ours)
    record_property("err_ref", e_ref)
    print("xbatched d6n4 merged: FP32 error %.3g (reference float %.3g)" % (e_ours, e_ref))
    assert e_ours <= RATIO * max(e_ref, U32)
    assert e_ours < 1e-5
