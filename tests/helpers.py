"""Shared test helpers: seeded inputs, reference closed-form inputs, checkers.

All randomness comes from splitmix64 (vectorised here, and in C in
oracle/kron_oracle.c; the two are pinned against each other), so any machine
regenerates exactly the inputs the fixtures were made from.
"""
import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(count, seed):
    """First `count` outputs of splitmix64 started at state `seed`."""
    with np.errstate(over="ignore"):
        s = np.uint64(seed) + _GOLD * np.arange(1, count + 1, dtype=np.uint64)
        z = (s ^ (s >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(count, seed, dtype=np.float64):
    """Uniform [-1, 1) from the top 53 bits of splitmix64."""
    k = (splitmix64(count, seed) >> np.uint64(11)).astype(np.float64)
    return (2.0 * (k * (1.0 / 9007199254740992.0)) - 1.0).astype(dtype)


def normal(count, seed, std=1.0, dtype=np.float64):
    """Box-Muller normal(0, std^2) from two splitmix64 streams."""
    u1 = (splitmix64(count, seed) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    u2 = (splitmix64(count, seed ^ 0x5DEECE66D) >> np.uint64(11)).astype(np.float64) * (
        1.0 / 9007199254740992.0)
    r = np.sqrt(-2.0 * np.log1p(-u1))
    return (std * r * np.cos(2.0 * np.pi * u2)).astype(dtype)


def integers(count, seed, lo, hi, dtype=np.float64):
    """Uniform integers in [lo, hi] (exact in any float type for small ranges)."""
    k = splitmix64(count, seed) >> np.uint64(11)
    return (lo + (k % np.uint64(hi - lo + 1)).astype(np.int64)).astype(dtype)


def sub_seed(seed, tag):
    return (int(seed) * 1000003 + int(tag) * 7919 + 12345) & 0xFFFFFFFFFFFFFFFF


# --- the reference tests' closed-form inputs ---------------------------------

def ref_x(size, batch, dtype=np.float64):
    """X(i,b) = (i + (b-1)*Xsize) / (Xsize*batch), 1-based
    (test_kronmult6_batched.cpp:148-161).  Returned as (batch, size)."""
    i = np.arange(1, size + 1, dtype=np.float64)[None, :]
    b = np.arange(1, batch + 1, dtype=np.float64)[:, None]
    return ((i + (b - 1) * size) / (size * batch)).astype(dtype)


def ref_a(n, d, batch, lda=None, dtype=np.float64, pad=np.nan):
    """A(i,j,k,b) = (i + (j-1)n + (k-1)n^2 + (b-1)*batch) / (n^2 d batch), 1-based
    (test_kronmult6_batched.cpp:165-175).  Returned as (batch, d, lda*n) with
    rows n+1..lda of each column set to `pad` (never read)."""
    lda = lda or n
    out = np.full((batch, d, n, lda), pad, dtype=np.float64)  # [b, k, j, i]
    i = np.arange(1, n + 1)[None, None, None, :]
    j = np.arange(1, n + 1)[None, None, :, None]
    k = np.arange(1, d + 1)[None, :, None, None]
    b = np.arange(1, batch + 1)[:, None, None, None]
    out[..., :n] = (i + (j - 1) * n + (k - 1) * n * n + (b - 1) * batch) / (n * n * d * batch)
    return out.reshape(batch, d, lda * n).astype(dtype)


# --- checkers ----------------------------------------------------------------

def factor_matrix(slab, m, n, ld):
    """m x n column-major factor stored in `slab` with leading dimension ld."""
    return np.asarray(slab, dtype=np.float64)[: ld * n].reshape(n, ld).T[:m, :]


def kron_numpy(factors, x):
    """numpy's kron chain: np.kron(A1, np.kron(A2, ...)) @ x (library routine)."""
    K = np.array([[1.0]])
    for A in factors:
        K = np.kron(K, A)
    return K @ np.asarray(x, dtype=np.float64)


def kron_modes(factors, x):
    """Mode-product evaluation (independent of the oracle's recursion):
    x is (batch, prod n) with the first factor on the slowest index."""
    x = np.asarray(x, dtype=np.float64)
    batch = x.shape[0]
    ns = [A.shape[1] for A in factors]
    t = x.reshape([batch] + ns)
    for k, A in enumerate(factors):
        t = np.moveaxis(np.tensordot(A, t, axes=([1], [k + 1])), 0, k + 1)
    return t.reshape(batch, -1)


def kron_modes_batched(Am, x):
    """Per-entry mode products: Am[b][k] is the m_k x n_k factor k of entry b,
    x is (batch, prod n).  Returns (batch, prod m)."""
    x = np.asarray(x, dtype=np.float64)
    batch = x.shape[0]
    d = len(Am[0])
    ns = [np.asarray(Am[0][k]).shape[1] for k in range(d)]
    t = x.reshape([batch] + ns)
    for k in range(d):
        Ak = np.stack([np.asarray(Am[b][k], dtype=np.float64) for b in range(batch)])
        tk = np.moveaxis(t, k + 1, 1)
        shp = tk.shape
        out = np.einsum("bij,bjr->bir", Ak, tk.reshape(batch, shp[1], -1))
        t = np.moveaxis(out.reshape((batch, Ak.shape[1]) + shp[2:]), 1, k + 1)
    return t.reshape(batch, -1)


def rel_err(got, want):
    """Reference tests' measure: max |got-want| / (1 + max(|got|,|want|))
    (test_kronmult6_xbatched.cpp:590-604)."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    if got.size == 0:
        return 0.0
    return float(np.max(np.abs(got - want) / (1.0 + np.maximum(np.abs(got), np.abs(want)))))


def norm_rel_err(got, want):
    """||got - want||_inf / ||want||_inf (0 when both are 0)."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    den = np.max(np.abs(want)) if want.size else 0.0
    num = np.max(np.abs(got - want)) if want.size else 0.0
    return float(num / den) if den > 0 else float(num)


def mixed_class_counts(total, seed, d_lo=2, d_hi=6, n_lo=3, n_hi=8, max_elems=65536):
    """Mixed-shape workload: `total` entries with d uniform in [d_lo, d_hi] and
    n uniform in [n_lo, n_hi], pairs with n^d > max_elems redrawn.  Returns
    {(d, n): count} in (d, n) order."""
    out = {}
    got = 0
    stream = 0
    while got < total:
        r = splitmix64(8 * total, sub_seed(seed, stream))
        ds = d_lo + (r[0::2] % np.uint64(d_hi - d_lo + 1)).astype(np.int64)
        ns = n_lo + (r[1::2] % np.uint64(n_hi - n_lo + 1)).astype(np.int64)
        keep = ns.astype(np.float64) ** ds <= max_elems
        for d, n in zip(ds[keep], ns[keep]):
            if got == total:
                break
            out[(int(d), int(n))] = out.get((int(d), int(n)), 0) + 1
            got += 1
        stream += 1
    return dict(sorted(out.items()))


def modes(F, t, skip=None):
    """Apply F[:, k] (batch, n, n) along mode k of t (batch, n, ..., n), k != skip."""
    d = F.shape[1]
    for k in range(d):
        if k == skip:
            continue
        tk = np.moveaxis(t, k + 1, 1)
        shp = tk.shape
        out = np.matmul(F[:, k], tk.reshape(shp[0], shp[1], -1))
        t = np.moveaxis(out.reshape(shp), 1, k + 1)
    return t


def grads_batched(d, n, A, X, dY):
    """Per-mode formula, vectorised over entries: dX = kron(A_k^T) dY;
    dA_k(p,q) = sum over the other indices of dY[..p..] * U_k[..q..], with
    U_k = X contracted by every factor except k.  A is (batch, d, n*n)
    column-major; dA is returned in the same layout."""
    batch = X.shape[0]
    F = np.swapaxes(A.reshape(batch, d, n, n).astype(np.float64), -1, -2)  # F[b,k][i,j]
    shape = (batch,) + (n,) * d
    dX = modes(np.swapaxes(F, -1, -2), dY.reshape(shape).astype(np.float64)).reshape(batch, -1)
    dA = np.zeros((batch, d, n * n))
    Yt = dY.reshape(shape).astype(np.float64)
    for k in range(d):
        U = modes(F, X.reshape(shape).astype(np.float64), skip=k)
        Yk = np.moveaxis(Yt, k + 1, 1).reshape(batch, n, -1)
        Uk = np.moveaxis(U, k + 1, 1).reshape(batch, n, -1)
        M = np.matmul(Yk, np.swapaxes(Uk, 1, 2))  # M[b][p, q]
        dA[:, k] = np.swapaxes(M, 1, 2).reshape(batch, -1)  # (p, q) at p + q*n
    return dX, dA
