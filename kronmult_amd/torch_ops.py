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

    Y = kron_xmatvec(Apool, amap, Xpool, xmap, ymap, ny)

The pointer-array layout (kronmult_xbatched.hpp:79-83) over pools: factor k
of entry b is Apool[amap[b, k]] (an n x n matrix), its X vector
Xpool[xmap[b]], and Y (ny, n^d) starts at zero with Y[ymap[b]] +=
kron(...) @ Xpool[xmap[b]].  Entries may share factors, X vectors and Y
targets.  Forward runs kron_xbatched_*, backward kron_xbatched_adjoint_*,
whose atomic adds sum the gradients of shared factors and X vectors.
"""
import torch

from . import (KRON_EWORKSPACE, KronError, adjoint_workspace_elems, kronmult_batched,
               kronmult_batched_adjoint, kronmult_xbatched, kronmult_xbatched_adjoint)


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
        try:
            kronmult_batched(d, n, Ac, Xc, Y, None, batch)
        except KronError as e:
            if e.status != KRON_EWORKSPACE:
                raise
            # entries too large for the chip stream through W (the call above
            # was refused before any launch)
            kronmult_batched(d, n, Ac, Xc, Y, torch.empty_like(Xc), batch)
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


def _rows(t, idx):
    """int64 device tensor of the addresses of rows t[idx] (t contiguous)."""
    step = t.stride(0) * t.element_size()
    return idx.reshape(-1).to(device=t.device, dtype=torch.int64) * step + t.data_ptr()


class KronXMatvec(torch.autograd.Function):
    @staticmethod
    def forward(ctx, Apool, Xpool, amap, xmap, ymap, ny):
        if Apool.dim() != 3 or Apool.shape[-1] != Apool.shape[-2]:
            raise ValueError("Apool must be (na, n, n)")
        if amap.dim() != 2:
            raise ValueError("amap must be (batch, d)")
        batch, d = amap.shape
        n = Apool.shape[-1]
        E = n ** d
        if Xpool.dim() != 2 or Xpool.shape[1] != E:
            raise ValueError("Xpool must be (nx, n^d)")
        if xmap.shape != (batch,) or ymap.shape != (batch,):
            raise ValueError("xmap and ymap must be (batch,)")
        if Apool.dtype != Xpool.dtype or Apool.device != Xpool.device:
            raise ValueError("Apool and Xpool must share dtype and device")
        for m, hi in ((amap, Apool.shape[0]), (xmap, Xpool.shape[0]), (ymap, ny)):
            if batch and (int(m.min()) < 0 or int(m.max()) >= hi):
                raise ValueError("map index out of range")
        Ac = _colmajor(Apool)
        Xc = Xpool.contiguous()
        Y = torch.zeros((ny, E), dtype=Xc.dtype, device=Xc.device)
        pA, pX, pY = _rows(Ac, amap), _rows(Xc, xmap), _rows(Y, ymap)
        try:
            kronmult_xbatched(d, n, pA, n, pX, pY, None, batch, dtype=Xc.dtype)
        except KronError as e:
            if e.status != KRON_EWORKSPACE:
                raise
            # entries too large for the chip: 2 * subbatch scratch vectors
            sub = min(batch, 1024)
            W = torch.empty((2 * sub, E), dtype=Xc.dtype, device=Xc.device)
            pW = _rows(W, torch.arange(2 * sub, device=Xc.device))
            kronmult_xbatched(d, n, pA, n, pX, pY, pW, batch, sub, dtype=Xc.dtype)
        ctx.save_for_backward(Ac, Xc, amap, xmap, ymap)
        ctx.dims = (batch, d, n)
        return Y

    @staticmethod
    def backward(ctx, dY):
        Ac, Xc, amap, xmap, ymap = ctx.saved_tensors
        batch, d, n = ctx.dims
        dYc = dY.contiguous()
        dX = torch.zeros_like(Xc)
        dA = torch.zeros_like(Ac)
        per = adjoint_workspace_elems(d, n, Xc.dtype)
        W = torch.empty(per * batch, dtype=Xc.dtype, device=Xc.device) if per else None
        kronmult_xbatched_adjoint(d, n, _rows(Ac, amap), n, _rows(Xc, xmap), _rows(dYc, ymap),
                                  _rows(dX, xmap), _rows(dA, amap), n, W, batch,
                                  dtype=Xc.dtype)
        return dA.transpose(-1, -2), dX, None, None, None, None


def kron_xmatvec(Apool, amap, Xpool, xmap, ymap, ny):
    """Y (ny, n^d) = 0, then Y[ymap[b]] += kron(Apool[amap[b, 0]], ...,
    Apool[amap[b, d-1]]) @ Xpool[xmap[b]] for every entry b; differentiable
    in Apool and Xpool."""
    return KronXMatvec.apply(Apool, Xpool, amap, xmap, ymap, ny)
