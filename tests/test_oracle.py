"""Pin the oracle (oracle/kron_oracle.c) without trusting its own formula.

Checks, each chosen so that a plausible mistake in the recursion (dropped
term, wrong stride, transposed factor, reversed mode order, wrong beta rule)
fails at least one of them:
  * numpy's np.kron chain on small shapes (library routine);
  * the C dense brute force (textbook definition, the reference tests' own
    checker) and an independent numpy mode-product evaluation;
  * exact integer arithmetic (bit-exact regardless of summation order);
  * permutation / identity / diagonal factors (closed forms);
  * the reference tests' closed-form inputs, sweeps and tolerances;
  * planted mistakes fed to the oracle must be caught by the same checks.
"""
import numpy as np
import pytest

import helpers as H
import oracle

O = oracle.oracle_lib()


def _batched_case(d, n, batch, seed, dtype=np.float64, kind="uniform"):
    E = n ** d
    if kind == "uniform":
        A = H.uniform(batch * d * n * n, H.sub_seed(seed, 1), dtype).reshape(batch, d, n * n)
        X = H.uniform(batch * E, H.sub_seed(seed, 2), dtype).reshape(batch, E)
    elif kind == "int":
        A = H.integers(batch * d * n * n, H.sub_seed(seed, 1), -2, 2, dtype).reshape(batch, d, n * n)
        X = H.integers(batch * E, H.sub_seed(seed, 2), -8, 8, dtype).reshape(batch, E)
    else:  # the reference's closed form
        A = H.ref_a(n, d, batch, dtype=dtype)
        X = H.ref_x(E, batch, dtype)
    return np.ascontiguousarray(A), np.ascontiguousarray(X)


def _run_batched(d, n, A, X, Y0=None):
    batch, E = X.shape
    Y = np.zeros_like(X) if Y0 is None else Y0.copy()
    Xc = X.copy()
    W = np.zeros_like(X)
    oracle.batched(O, d, n, A, Xc, Y, W, batch)
    return Y


def _factors(A, b, d, n, ld=None):
    ld = ld or n
    return [H.factor_matrix(A[b, k], n, n, ld) for k in range(d)]


@pytest.mark.parametrize("d", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("n", [1, 2, 3])
def test_batched_vs_numpy_kron(d, n):
    if n ** d > 800:
        pytest.skip("dense kron too large")
    A, X = _batched_case(d, n, 3, seed=d * 10 + n)
    Y = _run_batched(d, n, A, X)
    for b in range(3):
        want = H.kron_numpy(_factors(A, b, d, n), X[b])
        assert H.norm_rel_err(Y[b], want) < 1e-13


@pytest.mark.parametrize("d,n", [(6, 4), (5, 5), (4, 6), (3, 8), (2, 16), (1, 32)])
def test_batched_vs_dense_bruteforce(d, n):
    A, X = _batched_case(d, n, 2, seed=100 + d)
    Y = _run_batched(d, n, A, X)
    for b in range(2):
        f = _factors(A, b, d, n)
        want = oracle.dense_kron([n] * d, [n] * d, f, X[b])
        assert H.norm_rel_err(Y[b], want) < 1e-13
        assert H.norm_rel_err(Y[b], H.kron_modes(f, X[b:b + 1])[0]) < 1e-13


@pytest.mark.parametrize("d", [1, 2, 3, 4, 5, 6])
def test_reference_kat_sweep(d):
    """The reference sweep: batch {1,16,128} x n {1,2,3,4}, closed-form inputs,
    abs tolerance 1e-6 (test_kronmult6_batched.cpp:511-567), plus 1e-12 rel."""
    for batch in (1, 16, 128):
        for n in (1, 2, 3, 4):
            A, X = _batched_case(d, n, batch, 0, kind="kat")
            Y = _run_batched(d, n, A, X)
            want = H.kron_modes_batched([_factors(A, b, d, n) for b in range(batch)], X)
            assert np.max(np.abs(Y - want)) <= 1e-6
            assert H.rel_err(Y, want) < 1e-12


@pytest.mark.parametrize("d,n", [(6, 4), (4, 5), (3, 7), (2, 3), (1, 6)])
def test_integer_inputs_exact(d, n):
    A, X = _batched_case(d, n, 4, seed=7, kind="int")
    Y = _run_batched(d, n, A, X)
    for b in range(4):
        want = H.kron_modes(_factors(A, b, d, n), X[b:b + 1])[0]
        assert np.array_equal(Y[b], want)


@pytest.mark.parametrize("d,n", [(6, 4), (5, 3), (3, 5)])
def test_permutation_factors_exact(d, n):
    rng = np.random.default_rng(3)
    batch = 2
    A = np.zeros((batch, d, n * n))
    perms = rng.permuted(np.tile(np.arange(n), (batch, d, 1)), axis=2)
    for b in range(batch):
        for k in range(d):
            P = np.zeros((n, n))
            P[np.arange(n), perms[b, k]] = 1.0  # P(i, perm[i]) = 1 -> y_i = x_perm[i]
            A[b, k] = P.T.reshape(-1)           # column-major storage
    X = H.uniform(batch * n ** d, 5).reshape(batch, -1)
    Y = _run_batched(d, n, A, X)
    for b in range(batch):
        idx = np.indices([n] * d).reshape(d, -1)  # output multi-index, mode 1 slowest
        src = np.zeros(idx.shape[1], dtype=np.int64)
        for k in range(d):
            src = src * n + perms[b, k][idx[k]]
        assert np.array_equal(Y[b], X[b][src])


def test_identity_and_diagonal_closed_forms():
    d, n, batch = 5, 4, 3
    X = H.uniform(batch * n ** d, 9).reshape(batch, -1)
    A = np.tile(np.eye(n).reshape(-1), (batch, d, 1))
    Y0 = H.uniform(batch * n ** d, 10).reshape(batch, -1)
    assert np.array_equal(_run_batched(d, n, A, X, Y0), Y0 + X)
    diag = H.uniform(batch * d * n, 11).reshape(batch, d, n)
    A = np.zeros((batch, d, n * n))
    for b in range(batch):
        for k in range(d):
            A[b, k] = np.diag(diag[b, k]).reshape(-1)
    Y = _run_batched(d, n, A, X)
    for b in range(batch):
        scale = np.ones(1)
        for k in range(d):
            scale = np.multiply.outer(scale, diag[b, k]).reshape(-1)
        assert H.norm_rel_err(Y[b], scale * X[b]) < 1e-15


def test_linearity():
    d, n, batch = 4, 5, 2
    A, X1 = _batched_case(d, n, batch, 21)
    _, X2 = _batched_case(d, n, batch, 22)
    Y1, Y2 = _run_batched(d, n, A, X1), _run_batched(d, n, A, X2)
    Y12 = _run_batched(d, n, A, np.ascontiguousarray(1.5 * X1 - 2.0 * X2))
    assert H.norm_rel_err(Y12, 1.5 * Y1 - 2.0 * Y2) < 1e-13


def test_planted_mistakes_are_caught():
    """Mistakes planted in what the oracle is given must fail the checks."""
    d, n, batch = 4, 3, 2
    A, X = _batched_case(d, n, batch, 31)
    good = np.stack([H.kron_numpy(_factors(A, b, d, n), X[b]) for b in range(batch)])
    assert H.norm_rel_err(_run_batched(d, n, A, X), good) < 1e-13
    # transposed operand
    At = A.reshape(batch, d, n, n).transpose(0, 1, 3, 2).reshape(batch, d, n * n).copy()
    assert H.norm_rel_err(_run_batched(d, n, At, X), good) > 1e-3
    # reversed mode order
    Ar = np.ascontiguousarray(A[:, ::-1, :])
    assert H.norm_rel_err(_run_batched(d, n, Ar, X), good) > 1e-3
    # one factor dropped (replaced by identity)
    Ad = A.copy()
    Ad[:, 2] = np.eye(n).reshape(-1)
    assert H.norm_rel_err(_run_batched(d, n, Ad, X), good) > 1e-3


@pytest.mark.parametrize("d,n", [(6, 2), (4, 3), (3, 4), (1, 5)])
def test_xbatched_overlap_lda_padding(d, n):
    """lda = n+3 with NaN padding (never read), all entries add into Y(:,1)
    (test_kronmult6_xbatched.cpp:110,307-312), integer inputs -> exact."""
    batch, lda, E = 16, n + 3, n ** d
    A = H.integers(batch * d * n * n, 41, -2, 2).reshape(batch, d, n, n)
    A_pool = np.full((batch * d, n, lda), np.nan)
    A_pool[:, :, :n] = A.reshape(batch * d, n, n)
    A_pool = A_pool.reshape(batch * d, lda * n)
    X_pool = H.integers(batch * E, 42, -8, 8).reshape(batch, E)
    X_keep = X_pool.copy()
    Y_pool = np.zeros((1, E))
    W_pool = np.full((batch, E), np.nan)
    a_map = np.arange(batch * d).reshape(batch, d)
    oracle.xbatched(O, d, n, A_pool, a_map, lda, X_pool, np.arange(batch), Y_pool,
                    np.zeros(batch, dtype=int), W_pool, np.arange(batch), batch)
    want = np.zeros(E)
    for b in range(batch):
        f = [H.factor_matrix(A_pool[b * d + k], n, n, lda) for k in range(d)]
        want += H.kron_modes(f, X_keep[b:b + 1])[0]
    assert np.array_equal(Y_pool[0], want)
    assert np.array_equal(X_pool, X_keep)  # batch > 1: X copied, not clobbered


@pytest.mark.parametrize("seed", range(6))
def test_vbatched_rectangular_vs_dense(seed):
    rng = np.random.default_rng(seed)
    d = int(rng.integers(1, 7))
    m = [int(v) for v in rng.integers(1, 4, d)]
    n = [int(v) for v in rng.integers(1, 4, d)]
    batch = 3
    lds = m
    A_pool = np.zeros((batch * d, 9))
    for b in range(batch):
        for k in range(d):
            A_pool[b * d + k, : m[k] * n[k]] = H.uniform(m[k] * n[k], 1000 * seed + 10 * b + k)
    sx, sy = int(np.prod(n)), int(np.prod(m))
    X_pool = H.uniform(batch * sx, seed + 77).reshape(batch, sx)
    Y_pool = np.zeros((batch, sy))
    W = np.zeros(1 << 16)
    oracle.vbatched(O, m, n, A_pool, np.arange(batch * d).reshape(batch, d), X_pool,
                    np.arange(batch), Y_pool, np.arange(batch), W, W.nbytes, batch)
    for b in range(batch):
        f = [H.factor_matrix(A_pool[b * d + k], m[k], n[k], m[k]) for k in range(d)]
        want = oracle.dense_kron(m, n, f, X_pool[b])
        assert H.norm_rel_err(Y_pool[b], want) < 1e-13
        assert H.norm_rel_err(Y_pool[b], H.kron_modes(f, X_pool[b:b + 1])[0]) < 1e-13


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("beta", [0.0, 1.0, 1.2])
def test_kgemm_vs_numpy(trans, beta):
    """test_kgemm_nn_batched.cpp:85-86,202-219 inputs and beta rules."""
    for mm, nn, kk in [(1, 1, 1), (8, 15, 22), (29, 8, 1), (64, 36, 50)]:
        batch, alpha = 3, 1.3
        A = np.zeros((batch, mm * kk))
        B = np.zeros((batch, kk * nn))
        for b in range(batch):
            i, j = np.meshgrid(np.arange(1, mm + 1), np.arange(1, kk + 1), indexing="ij")
            A[b] = (1.0 + i + j * mm + b).T.reshape(-1)
            rb, cb = (kk, nn) if trans == "N" else (nn, kk)
            i, j = np.meshgrid(np.arange(1, rb + 1), np.arange(1, cb + 1), indexing="ij")
            B[b] = (1.0 / (1.0 + i + j * rb + b)).T.reshape(-1)
        C = np.full((batch, mm * nn), np.nan if beta == 0.0 else 1.0)
        ldb = kk if trans == "N" else nn
        oracle.kgemm(O, trans, mm, nn, kk, alpha, A, [mm] * batch, B, [ldb] * batch, beta, C,
                     [mm] * batch, batch)
        for b in range(batch):
            Am = A[b].reshape(kk, mm).T
            Bm = B[b].reshape(-1, ldb).T
            AB = Am @ (Bm if trans == "N" else Bm.T)
            c0 = 0.0 if beta == 0.0 else 1.0
            want = alpha * AB + beta * c0
            got = C[b].reshape(nn, mm).T
            assert np.max(np.abs(got - want) / (1 + np.abs(want))) < 1e-13
