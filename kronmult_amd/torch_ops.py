"""Optional torch.autograd wrapper over the C ABI (the drop-in API stays C++).

    Y = kron_matvec(A, X)

A: (batch, d, n, n) with A[b, k] the k-th factor of entry b as a matrix
   (row i, column j), X: (batch, n^d) with the reference's index convention
   (A1 acts on the slowest index).  Y: (batch, n^d), Y[b] = kron(A[b,0], ...,
   A[b,d-1]) @ X[b].  Forward runs kron_batched_*, backward the hand-written
   adjoint kron_batched_adjoint_* (dX and every dA_k).  Entries that fit on
   chip keep every intermediate there; larger ones (d = 6, n = 6 among them)
   stream two vectors per entry through a scratch W, allocated only then.
   FP64 and FP32, on the GPU only.
"""
import torch

from . import adjoint_workspace_elems, kronmult_batched, kronmult_batched_adjoint


def _colmajor(A):
    # factor (i, j) stored at i + j*n: the reference's (n, n, d, batch) layout
    return A.transpose(-1, -2).contiguous()


class KronMatvec(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, X):
        if A.dim() != 4 or A.shape[-1] != A.shape[-2]:
            raise ValueError("A must be (batch, d, n, n)")
        batch, d, n, _ = A.shape
        if X.shape != (batch, n ** d):
            raise ValueError("X must be (batch, n^d)")
        if A.dtype != X.dtype or A.device != X.device:
            raise ValueError("A and X must share dtype and device")
        Ac = _colmajor(A)
        Xc = X.contiguous()
        Y = torch.zeros_like(Xc)
        W = torch.empty_like(Xc)
        kronmult_batched(d, n, Ac, Xc, Y, W, batch)
        ctx.save_for_backward(Ac, Xc)
        ctx.dims = (batch, d, n)
        return Y

    @staticmethod
    def backward(ctx, dY):
        Ac, Xc = ctx.saved_tensors
        batch, d, n = ctx.dims
        dYc = dY.contiguous()
        dX = torch.empty_like(Xc)
        dA = torch.empty_like(Ac)
        per = adjoint_workspace_elems(d, n, Xc.dtype)
        W = torch.empty(per * batch, dtype=Xc.dtype, device=Xc.device) if per else None
        kronmult_batched_adjoint(d, n, Ac, Xc, dYc, dX, dA, W, batch)
        return dA.transpose(-1, -2), dX


def kron_matvec(A, X):
    """Y[b] = kron(A[b,0], ..., A[b,d-1]) @ X[b], differentiable in A and X."""
    return KronMatvec.apply(A, X)
