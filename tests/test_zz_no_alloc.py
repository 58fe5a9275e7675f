"""kron_xbatched_noalloc_f64 (the xbatched update with its sort scratch in a
fixed device array) through the C ABI, against kron_xbatched_f64 on the same
inputs: two calls in a row, the second smaller, so the second call runs on
whatever the first left in the array.  Integer-valued inputs: bit-exact;
uniform inputs: within 1e-12."""
import ctypes

import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import kronmult_amd as K  # noqa: E402
from gpu_util import DEV, dev, host, ptrs  # noqa: E402


def _pools(d, n, batch, nx, ny, seed, kind):
    E = n ** d
    if kind == "int":
        A = H.integers(batch * d * n * n, H.sub_seed(seed, 1), -2, 2)
        X = H.integers(nx * E, H.sub_seed(seed, 2), -8, 8)
    else:
        A = H.uniform(batch * d * n * n, H.sub_seed(seed, 1))
        X = H.uniform(nx * E, H.sub_seed(seed, 2))
    x_map = (H.splitmix64(batch, H.sub_seed(seed, 3)) % np.uint64(nx)).astype(np.int64)
    y_map = (H.splitmix64(batch, H.sub_seed(seed, 4)) % np.uint64(ny)).astype(np.int64)
    return A.reshape(batch * d, n * n), X.reshape(nx, E), x_map, y_map


def _run(fn, d, n, A, X, x_map, y_map, ny):
    batch = len(x_map)
    E = n ** d
    Ad, Xd = dev(A), dev(X)
    Yd = torch.zeros((ny, E), dtype=torch.float64, device=DEV)
    Wd = torch.empty((1, E), dtype=torch.float64, device=DEV)
    pA = ptrs(Ad, np.arange(batch * d))
    pX, pY = ptrs(Xd, x_map), ptrs(Yd, y_map)
    pW = ptrs(Wd, np.zeros(batch, dtype=np.int64))
    st = fn(d, n, pA.data_ptr(), n, pX.data_ptr(), pY.data_ptr(), pW.data_ptr(), batch, 0,
            torch.cuda.current_stream().cuda_stream)
    assert st == K.KRON_OK, K.lib().kron_status_string(st).decode()
    return host(Yd)


@pytest.mark.parametrize("kind", ["int", "uniform"])
def test_noalloc_matches_xbatched(kind):
    L = K.lib()
    cases = [(6, 4, 65536, 8192, 2048), (4, 5, 3001, 300, 97)]  # second call smaller
    for i, (d, n, batch, nx, ny) in enumerate(cases):
        A, X, x_map, y_map = _pools(d, n, batch, nx, ny, 40 + i, kind)
        got = _run(L.kron_xbatched_noalloc_f64, d, n, A, X, x_map, y_map, ny)
        want = _run(L.kron_xbatched_f64, d, n, A, X, x_map, y_map, ny)
        if kind == "int":
            assert np.array_equal(got, want), (d, n, batch)
        else:
            assert H.rel_err(got, want) < 1e-12, (d, n, batch)


def test_noalloc_rejects_above_limit():
    L = K.lib()
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: rejected on the host
    st = L.kron_xbatched_noalloc_f64(6, 4, fake, 4, fake, fake, fake,
                                     K.KRON_NOALLOC_MAX_BATCH + 1, 0, None)
    assert st == K.KRON_ETOOBIG
    assert L.kron_status_string(st).decode().startswith("KRON_ETOOBIG")
    # invalid arguments are still reported as by kron_xbatched_f64
    assert L.kron_xbatched_noalloc_f64(7, 4, fake, 4, fake, fake, fake, 10, 0, None) == \
        K.KRON_EINVAL
