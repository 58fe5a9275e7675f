"""The pointer-layout adjoint (kron_xbatched_adjoint_*) and its autograd
wrapper kron_xmatvec, over pools that entries share (factors, X vectors, Y
targets), against the per-mode gradient formula scattered into the pools,
the batched adjoint, central finite differences of the xbatched forward and
torch autograd of dense Kronecker products on the CPU.

Integer-valued inputs are bit-exact (every sum is exact in FP64, so the
order of the atomic adds does not matter); FP32 is held to the dot-product
error bound of its contractions and atomic sums, and to 1e-5 in norm."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import kronmult_amd as K  # noqa: E402
from kronmult_amd.torch_ops import kron_xmatvec  # noqa: E402
from gpu_util import DEV, dev, host, ptrs  # noqa: E402


def _maps(count, seed, hi):
    return (H.splitmix64(count, seed) % np.uint64(hi)).astype(np.int64)


def _pool_case(d, n, batch, na, nx, ny, lda, ldda, seed, integer=True, dtype=np.float64):
    """Pools and maps: factor k of entry b is pool row amap[b, k] (n x n,
    leading dimension lda); X through xmap, dY (the gradient of the entry's
    Y target) through ymap.  dX and dA pools start at nonzero values (the
    adjoint adds into them), padding included."""
    E = n ** d
    s = lambda t: H.sub_seed(seed, t)  # noqa: E731
    if integer:
        Ap = H.integers(na * lda * n, s(1), -2, 2).reshape(na, lda * n)
        Xp = H.integers(nx * E, s(2), -3, 3).reshape(nx, E)
        dYp = H.integers(ny * E, s(3), -3, 3).reshape(ny, E)
    else:
        Ap = H.uniform(na * lda * n, s(1)).reshape(na, lda * n)
        Xp = H.uniform(nx * E, s(2)).reshape(nx, E)
        dYp = H.uniform(ny * E, s(3)).reshape(ny, E)
    dX0 = H.integers(nx * E, s(4), -5, 5).reshape(nx, E)
    dA0 = H.integers(na * ldda * n, s(5), -5, 5).reshape(na, ldda * n)
    amap = _maps(batch * d, s(6), na).reshape(batch, d)
    xmap = _maps(batch, s(7), nx)
    ymap = _maps(batch, s(8), ny)
    return dict(d=d, n=n, lda=lda, ldda=ldda, Ap=Ap.astype(dtype), Xp=Xp.astype(dtype),
                dYp=dYp.astype(dtype), dX0=dX0.astype(dtype), dA0=dA0.astype(dtype), amap=amap,
                xmap=xmap, ymap=ymap)


def _entry_factors(c):
    """(batch, d, n*n) column-major factors of every entry (leading dim n)."""
    d, n, lda = c["d"], c["n"], c["lda"]
    batch = c["amap"].shape[0]
    F = c["Ap"][c["amap"].reshape(-1)].astype(np.float64).reshape(batch * d, n, lda)[:, :, :n]
    return F.reshape(batch, d, n * n)


def _expected(c):
    """Per-entry gradients (the per-mode formula) added into the pools."""
    d, n, ldda = c["d"], c["n"], c["ldda"]
    A = _entry_factors(c)
    dXe, dAe = H.grads_batched(d, n, A, c["Xp"][c["xmap"]], c["dYp"][c["ymap"]])
    dX = c["dX0"].astype(np.float64).copy()
    np.add.at(dX, c["xmap"], dXe)
    dA = c["dA0"].astype(np.float64).copy()
    # element (i, j) of factor row a sits at a*ldda*n + i + j*ldda
    i = np.arange(n)[None, None, None, :]
    j = np.arange(n)[None, None, :, None]
    flat = c["amap"][:, :, None, None] * (ldda * n) + j * ldda + i
    np.add.at(dA.reshape(-1), flat.reshape(-1), dAe.reshape(-1))
    return dX, dA


def _gpu(c, dtype=None):
    d, n = c["d"], c["n"]
    batch = c["amap"].shape[0]
    Ad, Xd, dYd = dev(c["Ap"]), dev(c["Xp"]), dev(c["dYp"])
    dXd, dAd = dev(c["dX0"]), dev(c["dA0"])
    X_before, dY_before, A_before = Xd.clone(), dYd.clone(), Ad.clone()
    per = K.adjoint_workspace_elems(d, n, Xd.dtype)
    Wd = torch.empty(per * batch, dtype=Xd.dtype, device=DEV) if per else None
    K.kronmult_xbatched_adjoint(d, n, ptrs(Ad, c["amap"]), c["lda"], ptrs(Xd, c["xmap"]),
                                ptrs(dYd, c["ymap"]), ptrs(dXd, c["xmap"]), ptrs(dAd, c["amap"]),
                                c["ldda"], Wd, batch, dtype=Xd.dtype)
    got = host(dXd), host(dAd)
    assert torch.equal(Xd, X_before) and torch.equal(dYd, dY_before) and torch.equal(Ad, A_before)
    return got


# every adjoint kernel family: compile-time n with 1..4 fast modes, the
# runtime-n kernel (n > 8), whole entries on chip, the d6n6 kernels
INT_CASES = [
    # d, n, batch, na, nx, ny
    (1, 5, 300, 7, 11, 5),
    (2, 3, 1000, 13, 17, 3),
    (3, 4, 501, 9, 40, 6),
    (4, 4, 4099, 31, 300, 50),
    (5, 3, 777, 15, 60, 9),
    (6, 4, 513, 40, 70, 20),
    (6, 4, 4099, 60, 300, 100),  # d6n4 on chip, entries in dX-target order
    (6, 4, 2048, 2048, 7, 2048),  # a few X vectors shared by ~290 entries each
    (4, 7, 120, 25, 30, 8),
    (5, 5, 97, 17, 21, 7),
    (3, 10, 64, 19, 23, 9),
    (2, 12, 300, 33, 37, 11),
    (6, 6, 150, 41, 37, 12),
]


@pytest.mark.parametrize("d,n,batch,na,nx,ny", INT_CASES)
def test_xadjoint_integer_bit_exact(d, n, batch, na, nx, ny):
    """Shared factors, X vectors and Y targets; padded leading dimensions
    (lda = n + 1, ldda = n + 2) whose padding must stay as it was."""
    c = _pool_case(d, n, batch, na, nx, ny, n + 1, n + 2, d * 100 + n)
    dX, dA = _gpu(c)
    wdX, wdA = _expected(c)
    assert np.array_equal(dX, wdX)
    assert np.array_equal(dA, wdA)


@pytest.mark.parametrize("d,n,batch", [(6, 6, 211), (4, 4, 2049), (3, 5, 3331), (6, 4, 3001)])
def test_xadjoint_matches_batched_adjoint(d, n, batch):
    """Private factors, X vectors and Y targets (the maps are the identity):
    the pointer layout adds what the batched layout overwrites (both sum the
    workgroups' shares of dA with atomics, in run-dependent order: FP64
    within 1e-12)."""
    E = n ** d
    A = H.uniform(batch * d * n * n, 301 + n).reshape(batch, d, n * n)
    X = H.uniform(batch * E, 302 + n).reshape(batch, E)
    dY = H.uniform(batch * E, 303 + n).reshape(batch, E)
    Ad, Xd, dYd = dev(A), dev(X), dev(dY)
    per = K.adjoint_workspace_elems(d, n, Xd.dtype)
    W = torch.empty(max(per, 1) * batch, dtype=Xd.dtype, device=DEV)
    dXb, dAb = torch.empty_like(Xd), torch.empty_like(Ad)
    K.kronmult_batched_adjoint(d, n, Ad, Xd, dYd, dXb, dAb, W if per else None, batch)
    A2 = Ad.reshape(batch * d, n * n)
    dXp, dAp = torch.zeros_like(Xd), torch.zeros_like(A2)
    ids = np.arange(batch)
    K.kronmult_xbatched_adjoint(d, n, ptrs(A2, np.arange(batch * d)), n, ptrs(Xd, ids),
                                ptrs(dYd, ids), ptrs(dXp, ids), ptrs(dAp, np.arange(batch * d)),
                                n, W if per else None, batch)
    assert H.rel_err(host(dXp), host(dXb)) < 1e-12
    assert H.rel_err(host(dAp).reshape(batch, d, -1), host(dAb)) < 1e-12


@pytest.mark.parametrize("d,n,batch", [(6, 4, 300), (6, 4, 2500), (6, 6, 40), (3, 8, 500)])
def test_xadjoint_f32(d, n, batch):
    c32 = _pool_case(d, n, batch, 23, 29, 7, n, n, 50 + n, integer=False, dtype=np.float32)
    c64 = dict(c32)
    for k in ("Ap", "Xp", "dYp", "dX0", "dA0"):
        c64[k] = c32[k].astype(np.float64)
    dX, dA = _gpu(c32)
    wdX, wdA = _expected(c64)
    # FP32 contractions (d of them, n terms each) and FP32 atomic sums of up
    # to `share` entries per pool row: the dot-product error bound
    # |err| <= k u sum|terms|, u = 2^-24, with sum|terms| the same gradients
    # computed from |A|, |X|, |dY| (and |dX0|, |dA0|)
    cabs = dict(c64)
    for k in ("Ap", "Xp", "dYp", "dX0", "dA0"):
        cabs[k] = np.abs(c64[k])
    bX, bA = _expected(cabs)
    share = max(np.bincount(c64["xmap"]).max(), np.bincount(c64["amap"].reshape(-1)).max())
    k = 2 * (d * n + share)
    u = 2.0 ** -24
    assert np.all(np.abs(dX - wdX) <= k * u * bX)
    assert np.all(np.abs(dA - wdA) <= k * u * bA)
    # and the usual FP32 bar in norm
    assert H.norm_rel_err(dX, wdX) < 1e-5
    assert H.norm_rel_err(dA, wdA) < 1e-5


def test_xadjoint_central_finite_differences():
    """L = sum over Y targets of <dY_t, Y_t>, Y from the xbatched forward on
    the GPU; central differences (step 1e-6) in sampled factor-pool and
    X-pool elements against the adjoint's dA and dX."""
    d, n, batch, h = 3, 4, 40, 1e-6
    c = _pool_case(d, n, batch, 9, 11, 5, n, n, 77, integer=False)
    c["dX0"] = np.zeros_like(c["dX0"])
    c["dA0"] = np.zeros_like(c["dA0"])
    dX, dA = _gpu(c)
    ny, E = c["dYp"].shape

    def L(Ap, Xp):
        Ad, Xd = dev(Ap), dev(Xp)
        Y = torch.zeros((ny, E), dtype=torch.float64, device=DEV)
        K.kronmult_xbatched(d, n, ptrs(Ad, c["amap"]), n, ptrs(Xd, c["xmap"]),
                            ptrs(Y, c["ymap"]), None, batch)
        return float((host(Y) * c["dYp"]).sum())

    rng = np.random.default_rng(3)
    for _ in range(10):
        a, e = int(rng.integers(c["Ap"].shape[0])), int(rng.integers(n * n))
        Ap, Am = c["Ap"].copy(), c["Ap"].copy()
        Ap[a, e] += h
        Am[a, e] -= h
        fd = (L(Ap, c["Xp"]) - L(Am, c["Xp"])) / (2 * h)
        assert abs(fd - dA[a, e]) < 1e-7 * (1 + abs(fd))
    for _ in range(10):
        x, e = int(rng.integers(c["Xp"].shape[0])), int(rng.integers(E))
        Xp, Xm = c["Xp"].copy(), c["Xp"].copy()
        Xp[x, e] += h
        Xm[x, e] -= h
        fd = (L(c["Ap"], Xp) - L(c["Ap"], Xm)) / (2 * h)
        assert abs(fd - dX[x, e]) < 1e-7 * (1 + abs(fd))


def test_xadjoint_rejects_bad_arguments():
    d, n = 3, 4
    t = torch.zeros(64, dtype=torch.float64, device=DEV)
    p = ptrs(t.reshape(1, -1), [0, 0, 0])
    with pytest.raises(K.KronError):  # ldda < n
        K.kronmult_xbatched_adjoint(d, n, p, n, p, p, p, p, n - 1, None, 1)
    with pytest.raises(K.KronError):  # lda < n
        K.kronmult_xbatched_adjoint(d, n, p, n - 1, p, p, p, p, n, None, 1)
    with pytest.raises(K.KronError):  # missing W where the entry needs scratch
        K.kronmult_xbatched_adjoint(6, 6, p, 6, p, p, p, p, 6, None, 1)
    # batchCount 0: nothing to do, no pointer is read
    K.kronmult_xbatched_adjoint(d, n, None, n, None, None, None, None, n, None, 0)


def test_kron_xmatvec_autograd_dense_cpu():
    """The wrapper against torch autograd of dense Kronecker products on the
    CPU, with shared factors, X vectors and Y targets."""
    d, n, batch, na, nx, ny = 3, 3, 30, 5, 6, 4
    Ap = torch.from_numpy(H.uniform(na * n * n, 91).reshape(na, n, n))
    Xp = torch.from_numpy(H.uniform(nx * n ** d, 92).reshape(nx, -1))
    G = torch.from_numpy(H.uniform(ny * n ** d, 93).reshape(ny, -1))
    amap = torch.from_numpy(_maps(batch * d, 94, na).reshape(batch, d))
    xmap = torch.from_numpy(_maps(batch, 95, nx))
    ymap = torch.from_numpy(_maps(batch, 96, ny))
    Ag = Ap.to(DEV).requires_grad_(True)
    Xg = Xp.to(DEV).requires_grad_(True)
    Y = kron_xmatvec(Ag, amap.to(DEV), Xg, xmap.to(DEV), ymap.to(DEV), ny)
    (Y * G.to(DEV)).sum().backward()
    Ac = Ap.clone().requires_grad_(True)
    Xc = Xp.clone().requires_grad_(True)
    Yc = torch.zeros((ny, n ** d), dtype=torch.float64)
    for b in range(batch):
        Kd = Ac[amap[b, 0]]
        for k in range(1, d):
            Kd = torch.kron(Kd, Ac[amap[b, k]])
        Yc = Yc.index_add(0, ymap[b:b + 1], (Kd @ Xc[xmap[b]])[None])
    (Yc * G).sum().backward()
    assert H.rel_err(host(Y.detach()), Yc.detach().numpy()) < 1e-13
    assert H.rel_err(host(Ag.grad), Ac.grad.numpy()) < 1e-13
    assert H.rel_err(host(Xg.grad), Xc.grad.numpy()) < 1e-13


def test_kron_xmatvec_large_entries():
    """d = 6, n = 6 through the wrapper (the forward and the adjoint both
    stream through scratch): gradients against the per-mode formula."""
    d, n, batch, na, nx, ny = 6, 6, 64, 20, 30, 9
    c = _pool_case(d, n, batch, na, nx, ny, n, n, 606)
    c["dX0"] = np.zeros_like(c["dX0"])
    c["dA0"] = np.zeros_like(c["dA0"])
    wdX, wdA = _expected(c)
    # the wrapper takes factors as (row, column) matrices
    Am = torch.from_numpy(c["Ap"].reshape(na, n, n)).transpose(-1, -2).contiguous()
    Ag = Am.to(DEV).requires_grad_(True)
    Xg = torch.from_numpy(c["Xp"]).to(DEV).requires_grad_(True)
    tmap = lambda m: torch.from_numpy(m).to(DEV)  # noqa: E731
    Y = kron_xmatvec(Ag, tmap(c["amap"]), Xg, tmap(c["xmap"]), tmap(c["ymap"]), ny)
    (Y * torch.from_numpy(c["dYp"]).to(DEV)).sum().backward()
    assert np.array_equal(host(Xg.grad), wdX)
    assert np.array_equal(host(Ag.grad.transpose(-1, -2).contiguous()).reshape(na, -1), wdA)


def test_kron_xmatvec_d6n4_pooled():
    """The headline shape through the wrapper, 2,500 entries over shared pools
    (the adjoint runs on chip, in dX-target order): integer data, bit-exact
    against the per-mode formula."""
    d, n, batch, na, nx, ny = 6, 4, 2500, 400, 300, 64
    c = _pool_case(d, n, batch, na, nx, ny, n, n, 646)
    c["dX0"] = np.zeros_like(c["dX0"])
    c["dA0"] = np.zeros_like(c["dA0"])
    wdX, wdA = _expected(c)
    Am = torch.from_numpy(c["Ap"].reshape(na, n, n)).transpose(-1, -2).contiguous()
    Ag = Am.to(DEV).requires_grad_(True)
    Xg = torch.from_numpy(c["Xp"]).to(DEV).requires_grad_(True)
    tmap = lambda m: torch.from_numpy(m).to(DEV)  # noqa: E731
    Y = kron_xmatvec(Ag, tmap(c["amap"]), Xg, tmap(c["xmap"]), tmap(c["ymap"]), ny)
    (Y * torch.from_numpy(c["dYp"]).to(DEV)).sum().backward()
    assert np.array_equal(host(Xg.grad), wdX)
    assert np.array_equal(host(Ag.grad.transpose(-1, -2).contiguous()).reshape(na, -1), wdA)
