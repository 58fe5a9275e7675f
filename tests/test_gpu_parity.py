"""GPU parity: the HIP library through its C ABI against the oracle.

Bars: FP64 within 1e-12 (the reference tests' relative measure,
|a-b| / (1 + max(|a|,|b|))); integer-valued inputs bit-exact, which pins
pointer and index handling regardless of atomic order; FP32 within the
reference's FP32 tolerance 1e-5 against the FP64 oracle.
"""
import numpy as np
import pytest

import helpers as H
import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import kronmult_amd as K  # noqa: E402
from gpu_util import DEV, dev, host, ints, ptrs  # noqa: E402

O = oracle.oracle_lib()
F64_TOL = 1e-12
F32_TOL = 1e-5


def _oracle_batched(d, n, A, X, Y0):
    Y, Xc, W = Y0.astype(np.float64), X.astype(np.float64), np.zeros(X.shape)
    oracle.batched(O, d, n, np.ascontiguousarray(A, dtype=np.float64), Xc, Y, W, X.shape[0])
    return Y


def _gpu_batched(d, n, A, X, Y0, with_w=True):
    Ad, Xd, Yd = dev(A), dev(X), dev(Y0)
    Wd = torch.empty_like(Xd) if with_w else None
    X_before = Xd.clone()
    K.kronmult_batched(d, n, Ad, Xd, Yd, Wd, X.shape[0])
    out = host(Yd)
    assert torch.equal(Xd, X_before), "X must not be written"
    return out


FAST_SHAPES = [(d, n) for d in range(1, 7) for n in range(1, 9) if n ** d <= 16807]


@pytest.mark.parametrize("d,n", FAST_SHAPES)
def test_batched_f64_uniform(d, n):
    batch = 7 if n ** d <= 4096 else 3
    E = n ** d
    A = H.uniform(batch * d * n * n, d * 100 + n).reshape(batch, d, n * n)
    X = H.uniform(batch * E, d * 100 + n + 50).reshape(batch, E)
    Y0 = H.uniform(batch * E, 3).reshape(batch, E)
    assert H.rel_err(_gpu_batched(d, n, A, X, Y0), _oracle_batched(d, n, A, X, Y0)) < F64_TOL


@pytest.mark.parametrize("d", range(1, 7))
def test_batched_reference_sweep(d):
    """test_kronmult6_batched.cpp:511-567: batch {1,16,128} x n {1,2,3,4},
    closed-form inputs."""
    for batch in (1, 16, 128):
        for n in (1, 2, 3, 4):
            E = n ** d
            A, X = H.ref_a(n, d, batch), H.ref_x(E, batch)
            Y0 = np.zeros((batch, E))
            got = _gpu_batched(d, n, A, X, Y0)
            want = _oracle_batched(d, n, A, X, Y0)
            assert np.max(np.abs(got - want)) <= 1e-6
            assert H.rel_err(got, want) < F64_TOL, (batch, n)


@pytest.mark.parametrize("d,n", [(6, 4), (5, 5), (3, 8), (4, 6), (2, 7), (1, 8)])
def test_batched_integer_bit_exact(d, n):
    batch, E = 33, n ** d
    A = H.integers(batch * d * n * n, 11, -2, 2).reshape(batch, d, n * n)
    X = H.integers(batch * E, 12, -8, 8).reshape(batch, E)
    Y0 = H.integers(batch * E, 13, -100, 100).reshape(batch, E)
    assert np.array_equal(_gpu_batched(d, n, A, X, Y0), _oracle_batched(d, n, A, X, Y0))


# (5, 8), (6, 5): entries whose LDS leaves one workgroup per CU (512 / 1,024
# threads per workgroup)
@pytest.mark.parametrize("d,n", [(6, 4), (4, 5), (3, 8), (2, 3), (1, 4), (5, 8), (6, 5)])
def test_batched_f32(d, n):
    batch, E = 9, n ** d
    A = H.uniform(batch * d * n * n, 21).reshape(batch, d, n * n)
    X = H.uniform(batch * E, 22).reshape(batch, E)
    Y0 = np.zeros((batch, E))
    got = _gpu_batched(d, n, A.astype(np.float32), X.astype(np.float32), Y0.astype(np.float32))
    want = _oracle_batched(d, n, A.astype(np.float32), X.astype(np.float32), Y0)
    assert H.rel_err(got, want) < F32_TOL


# generic kernel: n > 8 single pass, and entries larger than LDS (two passes
# through W)
@pytest.mark.parametrize("d,n,batch", [(2, 20, 5), (3, 12, 3), (6, 6, 3), (5, 8, 2),
                                       (3, 32, 2), (4, 16, 2), (1, 40, 4)])
def test_batched_generic_shapes(d, n, batch):
    E = n ** d
    A = H.uniform(batch * d * n * n, 31 + n).reshape(batch, d, n * n)
    X = H.uniform(batch * E, 32 + n).reshape(batch, E)
    Y0 = H.uniform(batch * E, 33).reshape(batch, E)
    assert H.rel_err(_gpu_batched(d, n, A, X, Y0), _oracle_batched(d, n, A, X, Y0)) < F64_TOL


def test_batched_errors():
    A = torch.zeros(16, dtype=torch.float64, device=DEV)
    with pytest.raises(K.KronError) as e:
        K.kronmult_batched(7, 2, A, A, A, A, 1)
    assert e.value.status == K.KRON_EINVAL
    with pytest.raises(K.KronError):
        K.kronmult_batched(2, 0, A, A, A, A, 1)
    # 6^6 needs the W scratch
    with pytest.raises(K.KronError) as e:
        K.kronmult_batched(6, 6, A, A, A, None, 1)
    assert e.value.status == K.KRON_EWORKSPACE
    K.kronmult_batched(3, 2, A, A, A, None, 0)  # empty batch is a no-op


# --- xbatched -----------------------------------------------------------------

def _xbatched_case(d, n, batch, nx, ny, lda, seed, kind="uniform", dtype=np.float64):
    E = n ** d
    gen = (lambda c, s: H.uniform(c, s)) if kind == "uniform" else (
        lambda c, s: H.integers(c, s, -2, 2))
    A = gen(batch * d * n * n, H.sub_seed(seed, 1)).reshape(batch * d, n, n)
    A_pool = np.full((batch * d, n, lda), np.nan)
    A_pool[:, :, :n] = A
    A_pool = A_pool.reshape(batch * d, lda * n).astype(dtype)
    xs = H.uniform(nx * E, H.sub_seed(seed, 2)) if kind == "uniform" else H.integers(
        nx * E, H.sub_seed(seed, 2), -8, 8)
    X_pool = xs.reshape(nx, E).astype(dtype)
    x_map = (H.splitmix64(batch, H.sub_seed(seed, 3)) % np.uint64(nx)).astype(np.int64)
    y_map = (H.splitmix64(batch, H.sub_seed(seed, 4)) % np.uint64(ny)).astype(np.int64)
    return A_pool, X_pool, x_map, y_map


def _oracle_xbatched(d, n, lda, A_pool, X_pool, x_map, Y0_pool, y_map):
    batch = len(x_map)
    Y = Y0_pool.astype(np.float64)
    W = np.zeros((max(2, batch), X_pool.shape[1]))
    oracle.xbatched(O, d, n, A_pool.astype(np.float64), np.arange(batch * d).reshape(batch, d),
                    lda, X_pool.astype(np.float64), x_map, Y, y_map, W, np.arange(W.shape[0]),
                    batch)
    return Y


def _gpu_xbatched(d, n, lda, A_pool, X_pool, x_map, Y0_pool, y_map, subbatch=0):
    batch = len(x_map)
    Ad, Xd, Yd = dev(A_pool), dev(X_pool), dev(Y0_pool)
    nslot = max(2, batch)
    Wd = torch.full((nslot, X_pool.shape[1]), float("nan"), dtype=Xd.dtype, device=DEV)
    pA = ptrs(Ad, np.arange(batch * d))
    pX, pY, pW = ptrs(Xd, x_map), ptrs(Yd, y_map), ptrs(Wd, np.arange(nslot))
    X_before = Xd.clone()
    K.kronmult_xbatched(d, n, pA, lda, pX, pY, pW, batch, subbatch, dtype=Xd.dtype)
    out = host(Yd)
    assert torch.equal(Xd, X_before), "X must not be written"
    return out


@pytest.mark.parametrize("d", range(1, 7))
def test_xbatched_reference_sweep(d):
    """test_kronmult6_xbatched.cpp: batch {1,16,128} x n {1,2,3,4}, lda = n+3,
    every entry adds into Y(:,1), closed-form inputs; tol 3e-7 there, 1e-12 here."""
    for batch in (1, 16, 128):
        for n in (1, 2, 3, 4):
            E, lda = n ** d, n + 3
            A_pool = H.ref_a(n, d, batch, lda=lda).reshape(batch * d, lda * n)
            X_pool = H.ref_x(E, batch)
            Y0 = np.zeros((1, E))
            y_map = np.zeros(batch, dtype=np.int64)
            got = _gpu_xbatched(d, n, lda, A_pool, X_pool, np.arange(batch), Y0, y_map)
            want = _oracle_xbatched(d, n, lda, A_pool, X_pool, np.arange(batch), Y0, y_map)
            assert H.rel_err(got, want) < F64_TOL, (batch, n)


@pytest.mark.parametrize("d,n", [(6, 4), (5, 3), (4, 5), (3, 8), (2, 6), (1, 7), (6, 5)])
def test_xbatched_shared_y_integer_bit_exact(d, n):
    batch, lda = 300, n + 2
    A_pool, X_pool, x_map, y_map = _xbatched_case(d, n, batch, 40, 9, lda, 5, kind="int")
    Y0 = np.zeros((9, n ** d))
    got = _gpu_xbatched(d, n, lda, A_pool, X_pool, x_map, Y0, y_map)
    want = _oracle_xbatched(d, n, lda, A_pool, X_pool, x_map, Y0, y_map)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("d,n", [(6, 4), (4, 7), (3, 12), (6, 6)])
def test_xbatched_uniform(d, n):
    batch, lda = 64 if n ** d < 10000 else 6, n
    A_pool, X_pool, x_map, y_map = _xbatched_case(d, n, batch, 16, 5, lda, 6)
    Y0 = H.uniform(5 * n ** d, 8).reshape(5, -1)
    got = _gpu_xbatched(d, n, lda, A_pool, X_pool, x_map, Y0, y_map, subbatch=3)
    want = _oracle_xbatched(d, n, lda, A_pool, X_pool, x_map, Y0, y_map)
    assert H.rel_err(got, want) < F64_TOL


def test_xbatched_f32():
    d, n, batch = 6, 4, 50
    A_pool, X_pool, x_map, y_map = _xbatched_case(d, n, batch, 10, 4, n, 9)
    Y0 = np.zeros((4, n ** d))
    got = _gpu_xbatched(d, n, n, A_pool.astype(np.float32), X_pool.astype(np.float32), x_map,
                        Y0.astype(np.float32), y_map)
    want = _oracle_xbatched(d, n, n, A_pool.astype(np.float32).astype(np.float64),
                            X_pool.astype(np.float32).astype(np.float64), x_map, Y0, y_map)
    assert H.rel_err(got, want) < F32_TOL


def test_xbatched_headline_shape_integer():
    """d=6, n=4, 65,536 entries, X pool 8,192, Y pool 2,048 (about 32 writers
    per Y): bit-exact with integer inputs at the full size."""
    d, n, batch = 6, 4, 65536
    A_pool, X_pool, x_map, y_map = _xbatched_case(d, n, batch, 8192, 2048, n, 77, kind="int")
    Y0 = np.zeros((2048, n ** d))
    got = _gpu_xbatched(d, n, n, A_pool, X_pool, x_map, Y0, y_map)
    want = _oracle_xbatched(d, n, n, A_pool, X_pool, x_map, Y0, y_map)
    assert np.array_equal(got, want)


# --- vbatched -----------------------------------------------------------------

def _vb_case(m, n, batch, seed, ny):
    d = len(m)
    A_pool = np.zeros((batch * d, max(mk * nk for mk, nk in zip(m, n))))
    for b in range(batch):
        for k in range(d):
            A_pool[b * d + k, : m[k] * n[k]] = H.uniform(m[k] * n[k], seed * 1000 + b * 10 + k)
    sx = int(np.prod(n))
    X_pool = H.uniform(batch * sx, seed + 5).reshape(batch, sx)
    y_map = np.arange(batch) % ny
    return A_pool, X_pool, y_map


def _sizes(m, n):
    d = len(m)
    sx = int(np.prod(n))
    sw = max(int(np.prod(m[:k + 1])) * int(np.prod(n[k + 1:])) for k in range(d))
    return sx, max(sx, sw)


@pytest.mark.parametrize("seed", range(12))
def test_vbatched_rectangular(seed):
    rng = np.random.default_rng(seed + 100)
    d = int(rng.integers(1, 7))
    hi = 4 if d > 3 else 9
    m = [int(v) for v in rng.integers(1, hi, d)]
    n = [int(v) for v in rng.integers(1, hi, d)]
    batch, ny = 10, 3
    A_pool, X_pool, y_map = _vb_case(m, n, batch, seed, ny)
    sy = int(np.prod(m))
    Y0 = np.zeros((ny, sy))
    _, sxw = _sizes(m, n)
    Wcap = 3 * 2 * sxw * 8
    Yw = Y0.copy()
    oracle.vbatched(O, m, n, A_pool, np.arange(batch * d).reshape(batch, d), X_pool,
                    np.arange(batch), Yw, y_map, np.zeros(Wcap // 8), Wcap, batch)
    Ad, Xd, Yd = dev(A_pool), dev(X_pool), dev(Y0)
    Wd = torch.empty(Wcap // 8, dtype=torch.float64, device=DEV)
    K.kronmult_vbatched(m, n, ptrs(Ad, np.arange(batch * d)), ptrs(Xd, np.arange(batch)),
                        ptrs(Yd, y_map), Wd, Wcap, batch)
    assert H.rel_err(host(Yd), Yw) < F64_TOL, (m, n)


# --- d = 1 beyond the fused kernels' n <= 8: the matrix-vector kernel -------

@pytest.mark.parametrize("n", [9, 16, 33, 64, 100])
def test_batched_d1_gemv(n):
    """One lane per output row; n = 100 takes two row steps per lane."""
    batch = 37
    A = H.integers(batch * n * n, 61, -2, 2).reshape(batch, 1, n * n)
    X = H.integers(batch * n, 62, -8, 8).reshape(batch, n)
    Y0 = H.integers(batch * n, 63, -9, 9).reshape(batch, n)
    assert np.array_equal(_gpu_batched(1, n, A, X, Y0), _oracle_batched(1, n, A, X, Y0))


@pytest.mark.parametrize("n,lda,batch", [(12, 12, 500), (12, 15, 500), (70, 71, 300),
                                         (9, 9, 150000)])
def test_xbatched_d1_gemv_shared_y(n, lda, batch):
    """Shared Y targets (atomic rows), a padded leading dimension, and a batch
    that takes several grid-stride steps."""
    ny = 7
    A_pool, X_pool, x_map, y_map = _xbatched_case(1, n, batch, 30, ny, lda, 9, kind="int")
    Y0 = np.zeros((ny, n))
    got = _gpu_xbatched(1, n, lda, A_pool, X_pool, x_map, Y0, y_map)
    want = _oracle_xbatched(1, n, lda, A_pool, X_pool, x_map, Y0, y_map)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("m,n", [(7, 13), (70, 5), (1, 40), (64, 64)])
def test_vbatched_d1_rectangular(m, n):
    batch, ny = 20, 3
    A_pool, X_pool, y_map = _vb_case([m], [n], batch, m * 100 + n, ny)
    Y0 = np.zeros((ny, m))
    Wcap = 2 * max(m, n) * 8
    Yw = Y0.copy()
    oracle.vbatched(O, [m], [n], A_pool, np.arange(batch).reshape(batch, 1), X_pool,
                    np.arange(batch), Yw, y_map, np.zeros(Wcap // 8), Wcap, batch)
    Ad, Xd, Yd = dev(A_pool), dev(X_pool), dev(Y0)
    Wd = torch.empty(Wcap // 8, dtype=torch.float64, device=DEV)
    K.kronmult_vbatched([m], [n], ptrs(Ad, np.arange(batch)), ptrs(Xd, np.arange(batch)),
                        ptrs(Yd, y_map), Wd, Wcap, batch)
    assert H.rel_err(host(Yd), Yw) < F64_TOL, (m, n)


# --- fused kernel with exactly n^(d-1) threads per entry (not a power of two) --

@pytest.mark.parametrize("d,n,batch", [(3, 6, 501), (5, 3, 250), (2, 5, 1003), (3, 3, 29),
                                       (2, 7, 37)])
def test_xbatched_exact_thread_groups(d, n, batch):
    """Several entries per workgroup with idle threads past the last one;
    batches that leave a partial last workgroup; shared Y (atomics)."""
    ny = 7
    A_pool, X_pool, x_map, y_map = _xbatched_case(d, n, batch, 11, ny, n, 41, kind="int")
    Y0 = np.zeros((ny, n ** d))
    got = _gpu_xbatched(d, n, n, A_pool, X_pool, x_map, Y0, y_map)
    want = _oracle_xbatched(d, n, n, A_pool, X_pool, x_map, Y0, y_map)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("d,n", [(3, 6), (5, 3)])
def test_batched_exact_thread_groups_f32(d, n):
    batch = 333
    E = n ** d
    A = H.uniform(batch * d * n * n, 121).reshape(batch, d, n * n)
    X = H.uniform(batch * E, 122).reshape(batch, E)
    Y0 = H.uniform(batch * E, 123).reshape(batch, E)
    got = _gpu_batched(d, n, A.astype(np.float32), X.astype(np.float32), Y0.astype(np.float32))
    want = _oracle_batched(d, n, A.astype(np.float32), X.astype(np.float32),
                           Y0.astype(np.float32))
    assert H.rel_err(got, want) < 8 * d * n * 2.0 ** -24


# --- pointer layout on the matrix cores: single-pass n = 16 / 32 shapes --------

@pytest.mark.parametrize("d,n,lda,batch,ny", [(3, 16, 18, 300, 7), (2, 32, 33, 200, 5),
                                              (2, 16, 16, 600, 600), (3, 16, 16, 5, 1)])
def test_xbatched_mfma_pointer_layout(d, n, lda, batch, ny):
    """Padded leading dimensions, Y shared by many entries (atomics) and
    nearly one Y per entry; integer values, so the comparison is exact."""
    A_pool, X_pool, x_map, y_map = _xbatched_case(d, n, batch, 20, ny, lda, 29, kind="int")
    Y0 = np.zeros((ny, n ** d))
    got = _gpu_xbatched(d, n, lda, A_pool, X_pool, x_map, Y0, y_map)
    want = _oracle_xbatched(d, n, lda, A_pool, X_pool, x_map, Y0, y_map)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("d,n,lda,batch,ny,subbatch", [(4, 16, 16, 7, 3, 2), (3, 32, 34, 5, 2, 0),
                                                       (5, 16, 16, 3, 2, 1), (4, 16, 17, 1, 1, 0)])
def test_xbatched_mfma_two_pass(d, n, lda, batch, ny, subbatch):
    """Two-pass n = 16 / 32 shapes in the pointer layout: the W slots
    pW[2*slot] reused group by group (subbatchCount slots), and a single entry
    (pW[0]); integer values, exact comparison."""
    A_pool, X_pool, x_map, y_map = _xbatched_case(d, n, batch, 4, ny, lda, 37, kind="int")
    Y0 = np.zeros((ny, n ** d))
    got = _gpu_xbatched(d, n, lda, A_pool, X_pool, x_map, Y0, y_map, subbatch=subbatch)
    want = _oracle_xbatched(d, n, lda, A_pool, X_pool, x_map, Y0, y_map)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("slots", [1, 2])
def test_vbatched_square_mfma_two_pass(slots):
    """Square vbatched d4n16 on the matrix cores through the W + slot*stride
    scratch rule, slot by slot."""
    m = n = [16] * 4
    batch, ny = 5, 2
    A_pool, X_pool, y_map = _vb_case(m, n, batch, 16, ny)
    _, sxw = _sizes(m, n)
    Y0 = np.zeros((ny, 16 ** 4))
    Yw = Y0.copy()
    Wcap = slots * 2 * sxw * 8
    oracle.vbatched(O, m, n, A_pool, np.arange(batch * 4).reshape(batch, 4), X_pool,
                    np.arange(batch), Yw, y_map, np.zeros(Wcap // 8), Wcap, batch)
    Ad, Xd, Yd = dev(A_pool), dev(X_pool), dev(Y0)
    Wd = torch.full((Wcap // 8,), float("nan"), dtype=torch.float64, device=DEV)
    K.kronmult_vbatched(m, n, ptrs(Ad, np.arange(batch * 4)), ptrs(Xd, np.arange(batch)),
                        ptrs(Yd, y_map), Wd, Wcap, batch)
    assert H.rel_err(host(Yd), Yw) < F64_TOL


@pytest.mark.parametrize("d,n", [(3, 16), (2, 32)])
def test_xbatched_mfma_pointer_layout_f32(d, n):
    batch, ny = 64, 9
    A_pool, X_pool, x_map, y_map = _xbatched_case(d, n, batch, 16, ny, n, 31)
    Y0 = np.zeros((ny, n ** d))
    got = _gpu_xbatched(d, n, n, A_pool.astype(np.float32), X_pool.astype(np.float32),
                        x_map, Y0.astype(np.float32), y_map)
    want = _oracle_xbatched(d, n, n, A_pool.astype(np.float32), X_pool.astype(np.float32),
                            x_map, Y0, y_map)
    # d n-term FP32 dot products per entry plus the atomic sums of up to
    # batch / ny entries: k = d*n + batch/ny roundings of 2^-24 (x8)
    assert H.rel_err(got, want) < 8 * (d * n + batch // ny + 1) * 2.0 ** -24


# --- square n = 9..15 at d >= 3: the generic kernel's row-block contraction --

@pytest.mark.parametrize("d,n", [(3, n) for n in range(9, 16)] +
                         [(4, n) for n in range(9, 16)] + [(5, 9)])
def test_batched_rowblocks(d, n):
    """d >= 4 covers every row-block geometry (two-row blocks, three-row
    blocks at n = 11, 12; a partial last block at n = 9, 11, 13, 15) in two
    passes through W; d = 3 the single-pass row blocks (n = 9, 12) and the
    runtime-extent kernel (the other n).  Integer values: exact comparison."""
    batch = 5 if d == 3 else 2
    E = n ** d
    A = H.integers(batch * d * n * n, 101 + n, -2, 2).reshape(batch, d, n * n)
    X = H.integers(batch * E, 102 + n, -4, 4).reshape(batch, E)
    Y0 = H.integers(batch * E, 103, -9, 9).reshape(batch, E)
    assert np.array_equal(_gpu_batched(d, n, A, X, Y0), _oracle_batched(d, n, A, X, Y0))


@pytest.mark.parametrize("d,n,lda,batch", [(3, 12, 14, 600), (3, 9, 9, 300), (3, 15, 15, 200),
                                           (4, 10, 11, 40)])
def test_xbatched_rowblocks_shared_y(d, n, lda, batch):
    """Shared Y targets (atomic adds), padded leading dimensions."""
    ny = 6
    A_pool, X_pool, x_map, y_map = _xbatched_case(d, n, batch, 25, ny, lda, 23, kind="int")
    Y0 = np.zeros((ny, n ** d))
    got = _gpu_xbatched(d, n, lda, A_pool, X_pool, x_map, Y0, y_map)
    want = _oracle_xbatched(d, n, lda, A_pool, X_pool, x_map, Y0, y_map)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("d,n", [(3, 10), (3, 11), (3, 14), (4, 12)])
def test_batched_rowblocks_f32(d, n):
    """FP32 takes the row blocks at every n = 9..15 (d = 3: single pass and,
    from n = 13, two passes)."""
    batch = 4
    E = n ** d
    A = H.uniform(batch * d * n * n, 111).reshape(batch, d, n * n)
    X = H.uniform(batch * E, 112).reshape(batch, E)
    Y0 = H.uniform(batch * E, 113).reshape(batch, E)
    got = _gpu_batched(d, n, A.astype(np.float32), X.astype(np.float32), Y0.astype(np.float32))
    want = _oracle_batched(d, n, A.astype(np.float32), X.astype(np.float32),
                           Y0.astype(np.float32))
    # d n-term FP32 dot products: k = d*n roundings of 2^-24 (x8 for partial sums)
    assert H.rel_err(got, want) < 8 * d * n * 2.0 ** -24


# --- d = 2 with extents up to 32 beyond the fused kernels: a wave per entry --

@pytest.mark.parametrize("n", [9, 12, 15, 20, 31])
def test_batched_d2_wave(n):
    batch = 41
    A = H.integers(batch * 2 * n * n, 71, -2, 2).reshape(batch, 2, n * n)
    X = H.integers(batch * n * n, 72, -8, 8).reshape(batch, n * n)
    Y0 = H.integers(batch * n * n, 73, -9, 9).reshape(batch, n * n)
    assert np.array_equal(_gpu_batched(2, n, A, X, Y0), _oracle_batched(2, n, A, X, Y0))


@pytest.mark.parametrize("n,lda,batch", [(12, 12, 3000), (10, 13, 700), (32, 32, 300),
                                         (16, 16, 500), (17, 17, 400), (24, 26, 300)])
def test_xbatched_d2_wave_shared_y(n, lda, batch):
    """n <= 16: a wave per entry (four slices of 32 KB at n = 16); from n = 17
    a workgroup per entry with its factors in LDS."""
    ny = 5
    A_pool, X_pool, x_map, y_map = _xbatched_case(2, n, batch, 40, ny, lda, 19, kind="int")
    Y0 = np.zeros((ny, n * n))
    got = _gpu_xbatched(2, n, lda, A_pool, X_pool, x_map, Y0, y_map)
    want = _oracle_xbatched(2, n, lda, A_pool, X_pool, x_map, Y0, y_map)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("m,n", [([7, 13], [11, 5]), ([3, 2], [5, 4]), ([32, 9], [10, 32]),
                                 ([1, 17], [20, 1])])
def test_vbatched_d2_rectangular(m, n):
    batch, ny = 23, 4
    A_pool, X_pool, y_map = _vb_case(m, n, batch, sum(m) * 10 + sum(n), ny)
    sy = m[0] * m[1]
    Y0 = np.zeros((ny, sy))
    _, sxw = _sizes(m, n)
    Wcap = 2 * sxw * 8
    Yw = Y0.copy()
    oracle.vbatched(O, m, n, A_pool, np.arange(batch * 2).reshape(batch, 2), X_pool,
                    np.arange(batch), Yw, y_map, np.zeros(Wcap // 8), Wcap, batch)
    Ad, Xd, Yd = dev(A_pool), dev(X_pool), dev(Y0)
    Wd = torch.empty(Wcap // 8, dtype=torch.float64, device=DEV)
    K.kronmult_vbatched(m, n, ptrs(Ad, np.arange(batch * 2)), ptrs(Xd, np.arange(batch)),
                        ptrs(Yd, y_map), Wd, Wcap, batch)
    assert H.rel_err(host(Yd), Yw) < F64_TOL, (m, n)


def test_batched_d2_wave_many_entries():
    """60,000 entries: several grid-stride steps of the wave-per-entry kernel."""
    n, batch = 10, 60000
    A = H.integers(batch * 2 * n * n, 81, -2, 2).reshape(batch, 2, n * n)
    X = H.integers(batch * n * n, 82, -4, 4).reshape(batch, n * n)
    Y0 = np.zeros((batch, n * n))
    assert np.array_equal(_gpu_batched(2, n, A, X, Y0), _oracle_batched(2, n, A, X, Y0))


@pytest.mark.parametrize("d,n,batch", [(1, 20, 50), (1, 65, 3), (1, 9, 1), (2, 12, 30),
                                       (2, 25, 7)])
def test_small_d_kernels_f32(d, n, batch):
    """FP32 through the d = 1 (matrix-vector) and d = 2 (wave per entry)
    kernels against the FP64 oracle on the same rounded inputs."""
    E = n ** d
    A = H.uniform(batch * d * n * n, 91).reshape(batch, d, n * n)
    X = H.uniform(batch * E, 92).reshape(batch, E)
    Y0 = H.uniform(batch * E, 93).reshape(batch, E)
    got = _gpu_batched(d, n, A.astype(np.float32), X.astype(np.float32), Y0.astype(np.float32))
    want = _oracle_batched(d, n, A.astype(np.float32), X.astype(np.float32),
                           Y0.astype(np.float32))
    # d n-term FP32 dot products: k = d*n roundings of 2^-24 (x8 for partial sums)
    assert H.rel_err(got, want) < 8 * d * n * 2.0 ** -24


def test_vbatched_large_two_pass_and_workspace_rule():
    m, n = [7, 6, 9, 8, 7], [8, 7, 6, 7, 8]
    batch = 3
    A_pool, X_pool, y_map = _vb_case(m, n, batch, 3, 2)
    sy = int(np.prod(m))
    _, sxw = _sizes(m, n)
    Y0 = np.zeros((2, sy))
    Yw = Y0.copy()
    Wcap = 2 * sxw * 8  # exactly one slot
    oracle.vbatched(O, m, n, A_pool, np.arange(batch * 5).reshape(batch, 5), X_pool,
                    np.arange(batch), Yw, y_map, np.zeros(Wcap // 8), Wcap, batch)
    Ad, Xd, Yd = dev(A_pool), dev(X_pool), dev(Y0)
    Wd = torch.empty(Wcap // 8, dtype=torch.float64, device=DEV)
    args = (ptrs(Ad, np.arange(batch * 5)), ptrs(Xd, np.arange(batch)), ptrs(Yd, y_map), Wd)
    K.kronmult_vbatched(m, n, *args, Wcap, batch)
    assert H.rel_err(host(Yd), Yw) < F64_TOL
    with pytest.raises(K.KronError) as e:
        K.kronmult_vbatched(m, n, *args, Wcap - 8, batch)
    assert e.value.status == K.KRON_EWORKSPACE


@pytest.mark.parametrize("nn,d,slots", [(12, 4, 2), (9, 3, 1), (15, 4, 3), (11, 4, 1)])
def test_vbatched_square_rowblocks(nn, d, slots):
    """Square vbatched entries at n = 9..15 take the row-block contraction;
    the workspace holds `slots` scratch slots, so two-pass shapes run in groups
    of that many entries."""
    m = n = [nn] * d
    batch, ny = 7, 3
    A_pool, X_pool, y_map = _vb_case(m, n, batch, nn, ny)
    _, sxw = _sizes(m, n)
    Y0 = np.zeros((ny, nn ** d))
    Yw = Y0.copy()
    Wcap = slots * 2 * sxw * 8
    oracle.vbatched(O, m, n, A_pool, np.arange(batch * d).reshape(batch, d), X_pool,
                    np.arange(batch), Yw, y_map, np.zeros(Wcap // 8), Wcap, batch)
    Ad, Xd, Yd = dev(A_pool), dev(X_pool), dev(Y0)
    Wd = torch.empty(Wcap // 8, dtype=torch.float64, device=DEV)
    K.kronmult_vbatched(m, n, ptrs(Ad, np.arange(batch * d)), ptrs(Xd, np.arange(batch)),
                        ptrs(Yd, y_map), Wd, Wcap, batch)
    assert H.rel_err(host(Yd), Yw) < F64_TOL, (nn, d)


@pytest.mark.parametrize("m,n", [([4, 5, 6], [4, 5, 6]), ([3, 2, 5, 4], [2, 5, 3, 3]),
                                 ([2, 3, 4, 2, 3, 4], [2, 3, 4, 2, 3, 4]),
                                 ([1, 7, 2], [6, 1, 8]), ([7, 7, 7, 6], [7, 6, 7, 7])])
def test_vbatched_small_entries_integer(m, n):
    """Mixed per-mode extents at d >= 3: a wave per entry while X, both
    intermediates and the factors fit 16 KB (all but the last shape, which
    takes the generic kernel); integer values, Y shared, exact comparison."""
    d = len(m)
    batch, ny = 301, 5
    rng = np.random.default_rng(sum(m) * 100 + sum(n))
    fs = max(a * b for a, b in zip(m, n))
    A_pool = np.zeros((batch * d, fs))
    for r in range(batch * d):
        k = r % d
        A_pool[r, :m[k] * n[k]] = rng.integers(-2, 3, m[k] * n[k])
    sx, sy = int(np.prod(n)), int(np.prod(m))
    X_pool = rng.integers(-3, 4, (batch, sx)).astype(np.float64)
    y_map = (np.arange(batch) * 7) % ny
    _, sxw = _sizes(m, n)
    Wcap = 2 * sxw * 8 * 4
    Y0 = np.zeros((ny, sy))
    Yw = Y0.copy()
    oracle.vbatched(O, m, n, A_pool, np.arange(batch * d).reshape(batch, d), X_pool,
                    np.arange(batch), Yw, y_map, np.zeros(Wcap // 8), Wcap, batch)
    Ad, Xd, Yd = dev(A_pool), dev(X_pool), dev(Y0)
    Wd = torch.full((Wcap // 8,), float("nan"), dtype=torch.float64, device=DEV)
    K.kronmult_vbatched(m, n, ptrs(Ad, np.arange(batch * d)), ptrs(Xd, np.arange(batch)),
                        ptrs(Yd, y_map), Wd, Wcap, batch)
    assert np.array_equal(host(Yd), Yw), (m, n)


def test_vbatched_small_entries_f32():
    m, n = [4, 5, 6], [3, 6, 5]
    d, batch, ny = 3, 97, 4
    A_pool, X_pool, y_map = _vb_case(m, n, batch, 77, ny)
    _, sxw = _sizes(m, n)
    Wcap = 2 * sxw * 4 * 2
    Y0 = np.zeros((ny, int(np.prod(m))))
    Yw = Y0.copy()
    A32, X32 = A_pool.astype(np.float32), X_pool.astype(np.float32)
    oracle.vbatched(O, m, n, A32.astype(np.float64), np.arange(batch * d).reshape(batch, d),
                    X32.astype(np.float64), np.arange(batch), Yw, y_map,
                    np.zeros(Wcap // 4), Wcap * 2, batch)
    Ad, Xd = dev(A32), dev(X32)
    Yd = torch.zeros((ny, int(np.prod(m))), dtype=torch.float32, device=DEV)
    Wd = torch.empty(Wcap // 4, dtype=torch.float32, device=DEV)
    K.kronmult_vbatched(m, n, ptrs(Ad, np.arange(batch * d)), ptrs(Xd, np.arange(batch)),
                        ptrs(Yd, y_map), Wd, Wcap, batch, dtype=torch.float32)
    # d dot products of at most 6 terms per output, plus the atomic sums of
    # about batch / ny entries: k = 18 + 25 roundings of 2^-24 (x8)
    assert H.rel_err(host(Yd), Yw) < 8 * (18 + batch // ny + 1) * 2.0 ** -24


def test_vbatched_square_headline_shape():
    m = n = [4] * 6
    batch = 100
    A_pool, X_pool, y_map = _vb_case(m, n, batch, 4, 7)
    Y0 = np.zeros((7, 4096))
    Yw = Y0.copy()
    Wcap = 2 * 4096 * 8
    oracle.vbatched(O, m, n, A_pool, np.arange(batch * 6).reshape(batch, 6), X_pool,
                    np.arange(batch), Yw, y_map, np.zeros(Wcap // 8), Wcap, batch)
    Ad, Xd, Yd = dev(A_pool), dev(X_pool), dev(Y0)
    Wd = torch.empty(Wcap // 8, dtype=torch.float64, device=DEV)
    K.kronmult6_vbatched(4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, ptrs(Ad, np.arange(batch * 6)),
                         ptrs(Xd, np.arange(batch)), ptrs(Yd, y_map), Wd, Wcap, batch)
    assert H.rel_err(host(Yd), Yw) < F64_TOL


# --- kgemm --------------------------------------------------------------------

@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("beta", [0.0, 1.0, 1.2])
def test_kgemm(trans, beta, dtype):
    """Subset of the reference sweep mm,nn,kk in 1..65 step 7 (test_kgemm_nn_batched.cpp
    :534-548) and its perf shapes; tolerance 1e-9 (f64) / 5e-4 (f32) there."""
    tol = 1e-12 if dtype == np.float64 else 1e-5
    for mm, nn, kk in [(1, 1, 1), (8, 15, 22), (64, 57, 43), (29, 1, 65), (4, 1024, 4),
                       (512, 8, 8)]:
        batch = 5
        A = H.uniform(batch * mm * kk, mm * 3 + kk, dtype).reshape(batch, -1)
        B = H.uniform(batch * nn * kk, nn + 5, dtype).reshape(batch, -1)
        C0 = (np.full((batch, mm * nn), np.nan) if beta == 0.0 else
              H.uniform(batch * mm * nn, 6).reshape(batch, -1)).astype(dtype)
        ldb = kk if trans == "N" else nn
        Cw = C0.copy()
        oracle.kgemm(O, trans, mm, nn, kk, 1.3, A, [mm] * batch, B, [ldb] * batch, beta, Cw,
                     [mm] * batch, batch)
        Ad, Bd, Cd = dev(A), dev(B), dev(C0)
        f = K.kgemm_nn_batched if trans == "N" else K.kgemm_nt_batched
        f(mm, nn, kk, 1.3, ptrs(Ad, np.arange(batch)), ints([mm] * batch),
          ptrs(Bd, np.arange(batch)), ints([ldb] * batch), beta, ptrs(Cd, np.arange(batch)),
          ints([mm] * batch), batch, dtype=Ad.dtype)
        got = host(Cd)
        assert np.all(np.isfinite(got))
        assert H.rel_err(got, Cw) < tol, (mm, nn, kk)


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("beta", [0.0, 1.0, 1.2])
@pytest.mark.parametrize("mm,nn,kk,batch", [(4, 4, 4, 1001), (5, 3, 7, 300), (1, 64, 3, 77),
                                            (16, 8, 2, 9)])
def test_kgemm_small_outputs_packed(mm, nn, kk, batch, trans, beta):
    """Outputs of at most 128 elements: several entries per workgroup, a
    partial last workgroup, padded leading dimensions (ldA = mm + 3,
    ldC = mm + 1)."""
    lda, ldc = mm + 3, mm + 1
    ldb = (kk if trans == "N" else nn) + 2
    bcols = nn if trans == "N" else kk
    A = np.full((batch, lda * kk), np.nan)
    B = np.full((batch, ldb * bcols), np.nan)
    C0 = np.full((batch, ldc * nn), 0.0 if beta == 0.0 else 0.5)
    for b in range(batch):
        a = H.uniform(mm * kk, 7 * b + 1).reshape(kk, mm)
        A[b].reshape(kk, lda)[:, :mm] = a
        rows = kk if trans == "N" else nn
        B[b].reshape(bcols, ldb)[:, :rows] = H.uniform(rows * bcols, 7 * b + 2).reshape(bcols, rows)
    Cw = C0.copy()
    oracle.kgemm(O, trans, mm, nn, kk, 1.3, np.nan_to_num(A), [lda] * batch, np.nan_to_num(B),
                 [ldb] * batch, beta, Cw, [ldc] * batch, batch)
    Ad, Bd, Cd = dev(A), dev(B), dev(C0)
    f = K.kgemm_nn_batched if trans == "N" else K.kgemm_nt_batched
    f(mm, nn, kk, 1.3, ptrs(Ad, np.arange(batch)), ints([lda] * batch),
      ptrs(Bd, np.arange(batch)), ints([ldb] * batch), beta, ptrs(Cd, np.arange(batch)),
      ints([ldc] * batch), batch)
    got = host(Cd)
    # the padding rows of C (i >= mm) are never written
    pad = got.reshape(batch, nn, ldc)[:, :, mm:]
    assert np.all(pad == (0.0 if beta == 0.0 else 0.5))
    assert H.rel_err(got, Cw) < 1e-12, (mm, nn, kk)


def _ref_kgemm_operand(nrow, ncol, ib, recip):
    """test_kgemm_nn_batched.cpp:195-213: entry ib's operand, 1-based (i, j),
    column-major with ld = nrow: 1 + i + j*nrow + ib, or its reciprocal."""
    i = np.arange(1, nrow + 1)[:, None]
    j = np.arange(1, ncol + 1)[None, :]
    v = 1.0 + i + j * nrow + ib
    return (1.0 / v if recip else v).T.reshape(-1)


@pytest.mark.parametrize("trans", ["N", "T"])
def test_kgemm_reference_sweep(trans):
    """The reference's whole sweep (test_kgemm_nn_batched.cpp:534-548 and
    test_kgemm_nt_batched.cpp:534-548): batchCount {1, 8, 15} x kk, nn, mm in
    1..65 step 7, its closed-form operands, C = 1, alpha = 1.3, beta = 1.2.
    It passes max_abserr < 1e-9 against its host product; here every entry
    of C is within 1e-12 (relative) of the oracle."""
    alpha, beta = 1.3, 1.2
    sizes = range(1, 66, 7)
    for batch in (1, 8, 15):
        for kk in sizes:
            for nn in sizes:
                for mm in sizes:
                    brow, bcol = (kk, nn) if trans == "N" else (nn, kk)
                    A = np.stack([_ref_kgemm_operand(mm, kk, b, False) for b in range(batch)])
                    B = np.stack([_ref_kgemm_operand(brow, bcol, b, True) for b in range(batch)])
                    C0 = np.ones((batch, mm * nn))
                    Cw = C0.copy()
                    oracle.kgemm(O, trans, mm, nn, kk, alpha, A, [mm] * batch, B,
                                 [brow] * batch, beta, Cw, [mm] * batch, batch)
                    Ad, Bd, Cd = dev(A), dev(B), dev(C0)
                    f = K.kgemm_nn_batched if trans == "N" else K.kgemm_nt_batched
                    f(mm, nn, kk, alpha, ptrs(Ad, np.arange(batch)), ints([mm] * batch),
                      ptrs(Bd, np.arange(batch)), ints([brow] * batch), beta,
                      ptrs(Cd, np.arange(batch)), ints([mm] * batch), batch)
                    got = host(Cd)
                    assert np.max(np.abs(got - Cw)) < 1e-9 * max(1.0, np.max(np.abs(Cw))), \
                        (batch, mm, nn, kk)
                    assert H.rel_err(got, Cw) < 1e-12, (batch, mm, nn, kk)


def test_kgemm_shared_c_atomic():
    """beta == 1 with every entry naming the same C: atomic accumulation."""
    mm, nn, kk, batch = 16, 16, 4, 200
    A = H.integers(batch * mm * kk, 1, -3, 3).reshape(batch, -1)
    B = H.integers(batch * kk * nn, 2, -3, 3).reshape(batch, -1)
    C0 = np.zeros((1, mm * nn))
    want = np.zeros((nn, mm))
    for b in range(batch):
        want += (A[b].reshape(kk, mm).T @ B[b].reshape(nn, kk).T).T
    Ad, Bd, Cd = dev(A), dev(B), dev(C0)
    K.kgemm_nn_batched(mm, nn, kk, 1.0, ptrs(Ad, np.arange(batch)), ints([mm] * batch),
                       ptrs(Bd, np.arange(batch)), ints([kk] * batch), 1.0,
                       ptrs(Cd, np.zeros(batch)), ints([mm] * batch), batch)
    assert np.array_equal(host(Cd)[0], want.reshape(-1))


# --- grouped (Y-merging) pointer-layout path: batch >= 2048 -------------------

# shapes with several entries per workgroup ((4, 4), (3, 4), (3, 8), (2, 8),
# (2, 2)) are not grouped: their per-element atomics cost less than the
# ordering's fixed ~20 us; they run here at the same batch sizes
@pytest.mark.parametrize("d,n", [(6, 4), (5, 4), (5, 5), (4, 5), (4, 4), (3, 4), (3, 8),
                                 (2, 8), (2, 2), (4, 6), (6, 3), (6, 5), (5, 6), (5, 7)])
@pytest.mark.parametrize("ny", [1, 3, 257, 0])
def test_xbatched_grouped_integer_bit_exact(d, n, ny):
    """Entries sorted by Y target and merged in registers; ny = 0 means every
    entry has its own Y (no sharing).  Integer inputs: exact."""
    batch = 2500 if n ** d <= 4096 else 2100
    ny_eff = batch if ny == 0 else ny
    A_pool, X_pool, x_map, y_map = _xbatched_case(d, n, batch, 97, ny_eff, n, 21 + ny,
                                                  kind="int")
    if ny == 0:
        y_map = np.random.default_rng(5).permutation(batch)
    Y0 = H.integers(ny_eff * n ** d, 3, -5, 5).reshape(ny_eff, -1)
    got = _gpu_xbatched(d, n, n, A_pool, X_pool, x_map, Y0, y_map)
    want = _oracle_xbatched(d, n, n, A_pool, X_pool, x_map, Y0, y_map)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("d,n", [(6, 4), (5, 8), (6, 5)])
def test_xbatched_grouped_uniform(dtype, d, n):
    """Uniform inputs at the grouping batch sizes.  FP64 d=5 n=8 exceeds one
    workgroup's LDS and runs the generic kernel (two passes, not grouped)."""
    batch, ny = 4096 if n ** d <= 4096 else 2100, 40
    A_pool, X_pool, x_map, y_map = _xbatched_case(d, n, batch, 300, ny, n, 8)
    Y0 = np.zeros((ny, n ** d))
    got = _gpu_xbatched(d, n, n, A_pool.astype(dtype), X_pool.astype(dtype), x_map,
                        Y0.astype(dtype), y_map)
    want = _oracle_xbatched(d, n, n, A_pool.astype(dtype).astype(np.float64),
                            X_pool.astype(dtype).astype(np.float64), x_map, Y0, y_map)
    # FP32: an output goes through d n-term contractions and the FP32 sum of
    # the batch/ny entries merged into its target, k = d*n + batch/ny
    # roundings of 2^-24 each; partial sums of mixed sign may exceed the
    # result, so the bound is 8 k 2^-24 (measured 1.3 / 2.2 / 4.3 k 2^-24 at
    # d6n4 / d6n5 / d5n8: 1e-5 / 1.1e-5 / 2.4e-5)
    k = d * n + batch / ny
    tol = F64_TOL if dtype == np.float64 else 8 * k * 2.0 ** -24
    assert H.rel_err(got, want) < tol


def test_xbatched_repeat_calls_accumulate():
    """Two calls on one stream reuse the library workspace; Y accumulates."""
    d, n, batch, ny = 6, 4, 3000, 11
    A_pool, X_pool, x_map, y_map = _xbatched_case(d, n, batch, 64, ny, n, 12, kind="int")
    Ad, Xd, Yd = dev(A_pool), dev(X_pool), dev(np.zeros((ny, n ** d)))
    pA = ptrs(Ad, np.arange(batch * d))
    pX, pY = ptrs(Xd, x_map), ptrs(Yd, y_map)
    for _ in range(2):
        K.kronmult6_xbatched(n, pA, n, pX, pY, pX, batch)
    want = _oracle_xbatched(d, n, n, A_pool, X_pool, x_map, np.zeros((ny, n ** d)), y_map)
    assert np.array_equal(host(Yd), 2 * want)


def test_xbatched_grouped_sizes_share_workspace():
    """Grouped calls of changing sizes on one stream share the library's
    ordering workspace: its bucket counts must be back at zero after every
    call, for any bucket count (2^11 .. 2^16), and a grown workspace starts
    cleared.  70,000 entries exceed one entry per counting thread."""
    d, n = 4, 5
    for batch, ny in [(2500, 7), (9000, 300), (2048, 1), (70000, 2000), (2600, 9),
                      (33000, 5000)]:
        A_pool, X_pool, x_map, y_map = _xbatched_case(d, n, batch, 97, ny, n, batch % 101,
                                                      kind="int")
        Y0 = H.integers(ny * n ** d, 3, -5, 5).reshape(ny, -1)
        got = _gpu_xbatched(d, n, n, A_pool, X_pool, x_map, Y0, y_map)
        want = _oracle_xbatched(d, n, n, A_pool, X_pool, x_map, Y0, y_map)
        assert np.array_equal(got, want), batch


def test_xbatched_grouped_long_chunks_by_counting():
    """150,000 entries at d=6 n=4: every workgroup walks ~147 positions, so its
    window of entry addresses (32 iterations) is refilled several times.  The
    entries combine 2 factor sets with 3 X vectors, so Y is known exactly from
    six oracle evaluations and the count of each (Y target, combination):
    an entry given another entry's X or factors changes those counts."""
    d, n, batch, ny = 6, 4, 150000, 64
    E = n ** d
    A_sets = H.integers(2 * d * n * n, 41, -2, 2).reshape(2 * d, n * n)
    X_pool = H.integers(3 * E, 42, -8, 8).reshape(3, E)
    a_of = (H.splitmix64(batch, 43) % np.uint64(2)).astype(np.int64)
    x_of = (H.splitmix64(batch, 44) % np.uint64(3)).astype(np.int64)
    y_of = (H.splitmix64(batch, 45) % np.uint64(ny)).astype(np.int64)
    Y0 = H.integers(ny * E, 46, -5, 5).reshape(ny, E)
    Ad, Xd, Yd = dev(A_sets), dev(X_pool), dev(Y0)
    pA = ptrs(Ad, (a_of[:, None] * d + np.arange(d)).reshape(-1))
    pX, pY = ptrs(Xd, x_of), ptrs(Yd, y_of)
    K.kronmult6_xbatched(n, pA, n, pX, pY, pX, batch)
    got = host(Yd)
    want = Y0.astype(np.float64)
    for a in range(2):
        for x in range(3):
            v = _oracle_xbatched(d, n, n, A_sets[a * d:(a + 1) * d], X_pool[x:x + 1],
                                 np.zeros(1, dtype=np.int64), np.zeros((1, E)),
                                 np.zeros(1, dtype=np.int64))[0]
            cnt = np.bincount(y_of[(a_of == a) & (x_of == x)], minlength=ny)
            want += cnt[:, None].astype(np.float64) * v[None, :]
    assert np.array_equal(got, want)


@pytest.mark.parametrize("d,n,batch,ny", [(6, 4, 3000, 7), (6, 4, 40, 3), (4, 5, 2500, 5),
                                          (6, 6, 5, 2)])
def test_xbatched_plan(d, n, batch, ny):
    """kron_xbatched_plan_*: order computed once, executed twice; X changed
    between executes (allowed), pointer arrays unchanged."""
    A_pool, X_pool, x_map, y_map = _xbatched_case(d, n, batch, 50, ny, n, 33, kind="int")
    Ad, Xd, Yd = dev(A_pool), dev(X_pool), dev(np.zeros((ny, n ** d)))
    Wd = torch.empty((max(2, batch), n ** d), dtype=torch.float64, device=DEV)
    pA, pX, pY = ptrs(Ad, np.arange(batch * d)), ptrs(Xd, x_map), ptrs(Yd, y_map)
    pW = ptrs(Wd, np.arange(max(2, batch)))
    plan = K.XbatchedPlan(d, n, pA, n, pX, pY, pW, batch)
    plan.execute()
    Xd.mul_(2.0)
    plan.execute()
    plan.close()
    want = _oracle_xbatched(d, n, n, A_pool, X_pool, x_map, np.zeros((ny, n ** d)), y_map)
    assert np.array_equal(host(Yd), 3 * want)


# --- pbatched ------------------------------------------------------------------

@pytest.mark.parametrize("d,n", [(1, 5), (2, 8), (3, 4), (4, 6), (5, 3), (6, 4), (6, 6),
                                 (3, 16), (4, 16)])
def test_pbatched_shared_y_integer_bit_exact(d, n):
    """kronmult{d}_pbatched: contiguous A, pointer X/Y/W, Y shared between
    entries (atomic last step), integer inputs, against the oracle; X must not
    be written; (6, 6) goes through the per-entry W vectors, (3, 16) through
    the matrix cores, (4, 16) through both."""
    batch, nx, ny = 40 if n ** d < 20000 else 6, 9, 5
    E = n ** d
    A = H.integers(batch * d * n * n, 31 + d, -2, 2).reshape(batch, d, n * n)
    X_pool = H.integers(nx * E, 32 + d, -3, 3).reshape(nx, E)
    Y0 = H.integers(ny * E, 33 + d, -5, 5).reshape(ny, E)
    x_map = (H.splitmix64(batch, 34 + d) % np.uint64(nx)).astype(np.int64)
    y_map = (H.splitmix64(batch, 35 + d) % np.uint64(ny)).astype(np.int64)
    want = Y0.copy()
    oracle.pbatched(O, d, n, A, X_pool, x_map, want, y_map, batch)
    Ad, Xd, Yd = dev(A), dev(X_pool), dev(Y0)
    Wd = torch.full((batch, E), float("nan"), dtype=torch.float64, device=DEV)
    X_before = Xd.clone()
    K.kronmult_pbatched(d, n, Ad, ptrs(Xd, x_map), ptrs(Yd, y_map), ptrs(Wd, np.arange(batch)),
                        batch)
    assert np.array_equal(host(Yd), want)
    assert torch.equal(Xd, X_before), "X must not be written"


def test_pbatched_f32_and_errors():
    d, n, batch = 4, 5, 33
    E = n ** d
    A = H.uniform(batch * d * n * n, 41).reshape(batch, d, n * n).astype(np.float32)
    X = H.uniform(batch * E, 42).reshape(batch, E).astype(np.float32)
    want = np.zeros((batch, E))
    oracle.pbatched(O, d, n, A.astype(np.float64), X.astype(np.float64), np.arange(batch), want,
                    np.arange(batch), batch)
    Ad, Xd = dev(A), dev(X)
    Yd = torch.zeros((batch, E), dtype=torch.float32, device=DEV)
    K.kronmult_pbatched(d, n, Ad, ptrs(Xd, np.arange(batch)), ptrs(Yd, np.arange(batch)), None,
                        batch)
    assert H.rel_err(host(Yd), want) < F32_TOL
    with pytest.raises(K.KronError) as e:
        K.kronmult_pbatched(6, 6, dev(np.zeros(6 * 36)), ptrs(Xd, [0]), ptrs(Yd, [0]), None, 1,
                            dtype=torch.float64)
    assert e.value.status == K.KRON_EWORKSPACE
    with pytest.raises(K.KronError) as e:
        K.kronmult_pbatched(7, 2, Ad, ptrs(Xd, [0]), ptrs(Yd, [0]), None, 1)
    assert e.value.status == K.KRON_EINVAL
