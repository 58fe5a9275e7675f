"""The C restatement against the reference CPU path compiled from its own
sources (oracle/_ref, built by oracle/build.py where the reference exists).

Same operation order -> bit-identical results wherever no two entries share
Y; with shared Y only the atomic order differs (integer inputs -> exact).
Skipped where oracle/_ref was not built.
"""
import numpy as np
import pytest

import helpers as H
import oracle

O = oracle.oracle_lib()
R = oracle.ref_lib()
pytestmark = pytest.mark.skipif(R is None, reason="oracle/_ref (reference build) not present")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("d", [1, 2, 3, 4, 5, 6])
def test_batched_bit_exact(d, dtype):
    for n in (1, 2, 3, 4, 5):
        if n ** d > 20000:
            continue
        batch, E = 5, n ** d
        A = H.uniform(batch * d * n * n, 100 + d * 10 + n, dtype).reshape(batch, d, n * n)
        X = H.uniform(batch * E, 200 + d * 10 + n, dtype).reshape(batch, E)
        Y0 = H.uniform(batch * E, 300 + n, dtype).reshape(batch, E)
        outs = []
        for L in (O, R):
            Xc, Y, W = X.copy(), Y0.copy(), np.zeros_like(X)
            oracle.batched(L, d, n, A, Xc, Y, W, batch)
            outs.append(Y)
        assert np.array_equal(outs[0], outs[1]), (d, n)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("d", [1, 3, 6])
def test_xbatched_bit_exact_disjoint_y(d, dtype):
    for n in (2, 3, 4):
        batch, lda, E = 9, n + 3, n ** d
        A_pool = H.uniform(batch * d * lda * n, 7 + n, dtype).reshape(batch * d, lda * n)
        X_pool = H.uniform(5 * E, 8 + n, dtype).reshape(5, E)
        x_map = np.arange(batch) % 5
        outs = []
        for L in (O, R):
            Y_pool = np.zeros((batch, E), dtype)
            W_pool = np.zeros((batch, E), dtype)
            oracle.xbatched(L, d, n, A_pool, np.arange(batch * d).reshape(batch, d), lda,
                            X_pool, x_map, Y_pool, np.arange(batch), W_pool, np.arange(batch),
                            batch, subbatch=3)
            outs.append(Y_pool)
        assert np.array_equal(outs[0], outs[1]), (d, n)


def test_xbatched_shared_y_integer_exact():
    d, n, batch, E = 5, 3, 40, 3 ** 5
    A_pool = H.integers(batch * d * n * n, 3, -2, 2).reshape(batch * d, n * n)
    X_pool = H.integers(batch * E, 4, -8, 8).reshape(batch, E)
    outs = []
    for L in (O, R):
        Y_pool = np.zeros((4, E))
        W_pool = np.zeros((batch, E))
        oracle.xbatched(L, d, n, A_pool, np.arange(batch * d).reshape(batch, d), n, X_pool,
                        np.arange(batch), Y_pool, np.arange(batch) % 4, W_pool,
                        np.arange(batch), batch)
        outs.append(Y_pool)
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("seed", range(8))
def test_vbatched_bit_exact(seed, dtype):
    rng = np.random.default_rng(seed)
    d = int(rng.integers(1, 7))
    m = [int(v) for v in rng.integers(1, 4, d)]
    n = [int(v) for v in rng.integers(1, 4, d)]
    batch = 5
    A_pool = H.uniform(batch * d * 9, seed, dtype).reshape(batch * d, 9)
    sx, sy = int(np.prod(n)), int(np.prod(m))
    X_pool = H.uniform(batch * sx, seed + 1, dtype).reshape(batch, sx)
    outs = []
    for L in (O, R):
        Y_pool = np.zeros((batch, sy), dtype)
        W = np.zeros(4096, dtype)
        # Wcapacity for two slots -> sub-batches of 2 (kronmult_vbatched.hpp:157)
        sizeXW = max(sx, max(int(np.prod(m[:k + 1])) * int(np.prod(n[k + 1:])) for k in range(d)))
        oracle.vbatched(L, m, n, A_pool, np.arange(batch * d).reshape(batch, d), X_pool,
                        np.arange(batch), Y_pool, np.arange(batch), W,
                        2 * 2 * sizeXW * W.itemsize, batch)
        outs.append(Y_pool)
    assert np.array_equal(outs[0], outs[1]), (m, n)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("beta", [0.0, 1.0, 1.2])
def test_kgemm_bit_exact(trans, beta, dtype):
    for mm, nn, kk in [(1, 1, 1), (8, 15, 22), (65, 64, 3), (29, 71, 8)]:
        batch = 3
        A = H.uniform(batch * mm * kk, mm + nn, dtype).reshape(batch, -1)
        B = H.uniform(batch * nn * kk, kk + 7, dtype).reshape(batch, -1)
        C0 = H.uniform(batch * mm * nn, 5, dtype).reshape(batch, -1)
        ldb = kk if trans == "N" else nn
        outs = []
        for L in (O, R):
            C = C0.copy()
            oracle.kgemm(L, trans, mm, nn, kk, 1.3, A, [mm] * batch, B, [ldb] * batch, beta, C,
                         [mm] * batch, batch)
            outs.append(C)
        assert np.array_equal(outs[0], outs[1]), (mm, nn, kk)


@pytest.mark.parametrize("d", [1, 2, 3, 5, 6])
def test_pbatched_disjoint_y_bit_exact_and_shared_y_integer(d):
    """kronmult{d}_pbatched (kronmult6_pbatched.hpp:14-81) in the reference
    against the restatement: the same per-entry recursion, so bit-identical
    with private Y; exact with integer inputs when Y is shared.  (d = 4: the
    reference's kronmult4_pbatched.cpp does not compile as shipped.)"""
    for n in (2, 3, 4):
        if n ** d > 8192:
            continue
        batch, E = 6, n ** d
        A = H.uniform(batch * d * n * n, 400 + d * 10 + n).reshape(batch, d, n * n)
        X = H.uniform(batch * E, 500 + d * 10 + n).reshape(batch, E)
        outs = []
        for L in (O, R):
            Y = np.zeros((batch, E))
            oracle.pbatched(L, d, n, A, X.copy(), np.arange(batch), Y, np.arange(batch), batch)
            outs.append(Y)
        assert np.array_equal(outs[0], outs[1]), (d, n)
        # shared Y, distinct X: the reference uses X as workspace for d >= 3
        # ("X is the input (but may be modified)", kronmult6_pbatched.hpp:5),
        # so entries sharing an X vector are not valid input there
        Ai = H.integers(batch * d * n * n, 600 + d, -2, 2).reshape(batch, d, n * n)
        Xi = H.integers(batch * E, 700 + d, -3, 3).reshape(batch, E)
        x_map, y_map = np.arange(batch), np.arange(batch) % 2
        outs = []
        for L in (O, R):
            Y = np.zeros((2, E))
            oracle.pbatched(L, d, n, Ai, Xi.copy(), x_map, Y, y_map, batch)
            outs.append(Y)
        assert np.array_equal(outs[0], outs[1]), (d, n)
