"""GPU-side helpers for the parity tests (torch is only device memory here)."""
import numpy as np
import torch

DEV = "cuda:0"


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def ptrs(pool, idx):
    """int64 device tensor of addresses of rows pool[idx] (pool: 2-D device tensor)."""
    base = pool.data_ptr()
    row = pool.stride(0) * pool.element_size()
    idx = np.asarray(idx, dtype=np.int64).reshape(-1)
    return torch.from_numpy(base + idx * row).to(DEV)


def ints(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).to(DEV)
