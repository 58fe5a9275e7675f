"""The hand-written adjoint (kron_batched_adjoint_*) and the torch autograd
wrapper, against dense-Kronecker gradients, an independent per-mode formula,
central finite differences and the inner-product identity."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import kronmult_amd as K  # noqa: E402
from kronmult_amd.torch_ops import kron_matvec  # noqa: E402
from gpu_util import DEV, dev, host  # noqa: E402


def _case(d, n, batch, seed, dtype=np.float64):
    A = H.uniform(batch * d * n * n, H.sub_seed(seed, 1)).reshape(batch, d, n * n).astype(dtype)
    X = H.uniform(batch * n ** d, H.sub_seed(seed, 2)).reshape(batch, -1).astype(dtype)
    dY = H.uniform(batch * n ** d, H.sub_seed(seed, 3)).reshape(batch, -1).astype(dtype)
    return A, X, dY


def _gpu_adjoint(d, n, A, X, dY):
    batch, E = X.shape
    Ad, Xd, dYd = dev(A), dev(X), dev(dY)
    dXd = torch.full_like(Xd, float("nan"))
    dAd = torch.full_like(Ad, float("nan"))
    Wd = torch.empty(2 * E * batch, dtype=Xd.dtype, device=DEV)
    K.kronmult_batched_adjoint(d, n, Ad, Xd, dYd, dXd, dAd, Wd, batch)
    return host(dXd), host(dAd)


def _factors(A, b, d, n):
    return [H.factor_matrix(A[b, k], n, n, n) for k in range(d)]


def _reference_grads(d, n, A, X, dY):
    """Independent per-mode formula: dX = kron(A)^T dY (mode products with the
    transposes); dA_k(p,q) = sum dY[.., p, ..] * U_k[.., q, ..] with U_k = X
    contracted by every factor except k."""
    batch = X.shape[0]
    dX = np.zeros_like(X, dtype=np.float64)
    dA = np.zeros((batch, d, n * n))
    for b in range(batch):
        f = _factors(A, b, d, n)
        dX[b] = H.kron_modes([F.T for F in f], dY[b:b + 1])[0]
        for k in range(d):
            g = list(f)
            g[k] = np.eye(n)
            Uk = H.kron_modes(g, X[b:b + 1])[0].reshape([n] * d)
            Yk = dY[b].reshape([n] * d)
            M = np.tensordot(np.moveaxis(Yk, k, 0), np.moveaxis(Uk, k, 0),
                             axes=(list(range(1, d)), list(range(1, d))))
            dA[b, k] = M.T.reshape(-1)  # column-major (p + q*n)
    return dX, dA


@pytest.mark.parametrize("d,n", [(1, 3), (1, 8), (2, 4), (3, 4), (3, 5), (4, 3), (4, 4),
                                 (5, 3), (5, 4), (6, 2), (6, 3), (6, 4)])
def test_adjoint_vs_mode_formula(d, n):
    A, X, dY = _case(d, n, 3, d * 10 + n)
    dX, dA = _gpu_adjoint(d, n, A, X, dY)
    rdX, rdA = _reference_grads(d, n, A, X, dY)
    assert H.rel_err(dX, rdX) < 1e-12
    assert H.rel_err(dA, rdA) < 1e-12


@pytest.mark.parametrize("d,n", [(6, 6), (5, 7), (4, 9), (3, 16)])
def test_adjoint_large_entries_streaming(d, n):
    """Entries larger than LDS: V and U stream through W."""
    A, X, dY = _case(d, n, 2, 70 + n)
    dX, dA = _gpu_adjoint(d, n, A, X, dY)
    rdX, rdA = _reference_grads(d, n, A, X, dY)
    assert H.rel_err(dX, rdX) < 1e-12
    assert H.rel_err(dA, rdA) < 1e-12


def test_adjoint_dense_kron_torch_cpu():
    """Dense kron gradients from torch autograd on the CPU (float64)."""
    d, n, batch = 3, 4, 8
    A, X, dY = _case(d, n, batch, 5)
    dX, dA = _gpu_adjoint(d, n, A, X, dY)
    for b in range(batch):
        fs = [torch.tensor(F, requires_grad=True) for F in _factors(A, b, d, n)]
        x = torch.tensor(X[b], requires_grad=True)
        Kd = fs[0]
        for F in fs[1:]:
            Kd = torch.kron(Kd, F)
        L = torch.dot(torch.tensor(dY[b]), Kd @ x)
        L.backward()
        assert np.max(np.abs(x.grad.numpy() - dX[b])) < 1e-12
        for k in range(d):
            want = fs[k].grad.numpy().T.reshape(-1)  # column-major
            assert np.max(np.abs(want - dA[b, k])) < 1e-12


def test_adjoint_central_finite_differences():
    """d=3, n=4, batch 8, central step 1e-6 on L = <dY, kron(A) X>, the
    forward evaluated by the GPU kernel."""
    d, n, batch, h = 3, 4, 8, 1e-6
    A, X, dY = _case(d, n, batch, 9)
    dX, dA = _gpu_adjoint(d, n, A, X, dY)

    def L(Am, Xm):
        Y = torch.zeros(batch * n ** d, dtype=torch.float64, device=DEV)
        K.kronmult_batched(d, n, dev(Am), dev(Xm), Y, None, batch)
        return (host(Y).reshape(batch, -1) * dY).sum(axis=1)

    rng = np.random.default_rng(0)
    for _ in range(12):
        k, idx = int(rng.integers(d)), int(rng.integers(n * n))
        Ap, Am = A.copy(), A.copy()
        Ap[:, k, idx] += h
        Am[:, k, idx] -= h
        fd = (L(Ap, X) - L(Am, X)) / (2 * h)
        assert np.max(np.abs(fd - dA[:, k, idx])) < 1e-7 * (1 + np.max(np.abs(fd)))
    for _ in range(12):
        j = int(rng.integers(n ** d))
        Xp, Xm = X.copy(), X.copy()
        Xp[:, j] += h
        Xm[:, j] -= h
        fd = (L(A, Xp) - L(A, Xm)) / (2 * h)
        assert np.max(np.abs(fd - dX[:, j])) < 1e-7 * (1 + np.max(np.abs(fd)))


def test_adjoint_inner_product_identity():
    d, n, batch = 6, 4, 16
    A, X, dY = _case(d, n, batch, 13)
    dX, _ = _gpu_adjoint(d, n, A, X, dY)
    Y = torch.zeros(batch * n ** d, dtype=torch.float64, device=DEV)
    K.kronmult_batched(d, n, dev(A), dev(X), Y, None, batch)
    lhs = (host(Y).reshape(batch, -1) * dY).sum(axis=1)
    rhs = (dX * X).sum(axis=1)
    assert np.max(np.abs(lhs - rhs) / (1 + np.abs(lhs))) < 1e-12


def test_adjoint_f32():
    d, n, batch = 6, 4, 4
    A, X, dY = _case(d, n, batch, 17)
    dX, dA = _gpu_adjoint(d, n, A.astype(np.float32), X.astype(np.float32),
                          dY.astype(np.float32))
    rdX, rdA = _reference_grads(d, n, A.astype(np.float32).astype(np.float64),
                                X.astype(np.float32).astype(np.float64),
                                dY.astype(np.float32).astype(np.float64))
    assert H.rel_err(dX, rdX) < 1e-5
    assert H.rel_err(dA, rdA) < 1e-5


def test_autograd_wrapper():
    d, n, batch = 4, 3, 5
    A, X, dY = _case(d, n, batch, 21)
    Am = torch.tensor(A.reshape(batch, d, n, n)).transpose(-1, -2).contiguous()  # [b,k,i,j]
    Ag = Am.to(DEV).requires_grad_(True)
    Xg = torch.tensor(X).to(DEV).requires_grad_(True)
    Y = kron_matvec(Ag, Xg)
    (Y * torch.tensor(dY).to(DEV)).sum().backward()
    # CPU dense reference
    Ac = Am.clone().requires_grad_(True)
    Xc = torch.tensor(X).requires_grad_(True)
    outs = []
    for b in range(batch):
        Kd = Ac[b, 0]
        for k in range(1, d):
            Kd = torch.kron(Kd, Ac[b, k])
        outs.append(Kd @ Xc[b])
    Yc = torch.stack(outs)
    (Yc * torch.tensor(dY)).sum().backward()
    assert torch.allclose(Y.detach().cpu(), Yc.detach(), rtol=0, atol=1e-12)
    assert torch.allclose(Ag.grad.cpu(), Ac.grad, rtol=0, atol=1e-12)
    assert torch.allclose(Xg.grad.cpu(), Xc.grad, rtol=0, atol=1e-12)


def test_adjoint_workspace_rule():
    A = torch.zeros(6 * 36, dtype=torch.float64, device=DEV)
    V = torch.zeros(6 ** 6, dtype=torch.float64, device=DEV)
    with pytest.raises(K.KronError) as e:
        K.kronmult_batched_adjoint(6, 6, A, V, V, V, A, None, 1)
    assert e.value.status == K.KRON_EWORKSPACE
